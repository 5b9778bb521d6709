"""Writer for IrisFDTD-Academic ``inputFDTD.dat`` files.

The layout (section headers and the order of the positional values) is the
one the reference parser reads (reference code/files.f90:39-344, example
inputs example_1/inputFDTD.dat).  Workloads are described as keyword
arguments and rendered to the text the reference and this engine both read.
"""

DEFAULTS = dict(
    description="iris-mi355x workload",
    i_pol=2, q=5.0, ixmax=16, iymax=16, izmax=1024, cfl=0.5, wzpml=0.09, npml=2, pmlzi=20, pmlzf=1005,
    plane_wave=1, tincr=20.0, tdecay=180.0, epump=1.0, wlpump=700.0, pump_dt=4.0, def_pulse=4, tfsf_o=30, cosine=0,
    no_downloads=10, output_time_step=1000, texit_fs=None, nt=8000, hi=503, sdl=(0, 0, 0),
    eps1=1.0, eps_bbox=2.25, eps3=1.0,
    box=(1, 1, 1, 1, 25, 1000), mbox=None,
    nof=200, dlambda=600.0 / 199.0, lno=400.0, to_fourier=0, sample_fourier=1,
    dh=20, dx=0, precision=None,
    sdl_opts=None,
)

# Snapshot options ("Time evolution options", reference code/sdl.f90:868-897):
# per axis X, Y, Z the ranges (xi, xf, yi, yf, zi, zf) -- only the ones of
# enabled planes matter -- the modulus flag E (1: |E|, 2: |H|) and the six
# component modes (Ex..Hz; 0 off, 1 Re, 2 Im, 3 |F|, 4 |Re|, 5 |Im|).
# layers = ((nX, nY, nZ), [dX...], [dY...], [dZ...]) writes the optional
# "Extra parameters-SDLextended" section (reference code/files.f90:291-319).
SDL_DEFAULTS = dict(copy=(0, 1), xi=(1, 1, 1), xf=(1, 1, 1), yi=(1, 1, 1), yf=(1, 1, 1), zi=(1, 1, 1),
                    zf=(1, 1, 1), module_e=1, fields=(0, 0, 0, 0, 0, 0), layers=None)

# dt(fs) of a run: CFL/(c0*sqrt(3/Q^2)) * uat with the reference's constants;
# used to turn a step count into the itmax (fs) entry so that
# int(texit_fs/dt_fs) reproduces the step count.
_C0 = 137.035995483398438
_F_ABOHR = 18.8972606658935547
_UAT = 0.0241888426244258881


def dt_fs(q_nm, cfl):
    q = _F_ABOHR * q_nm
    return cfl / (_C0 * (3.0 / q ** 2) ** 0.5) * _UAT


def render(**kw):
    """inputFDTD.dat text for the given overrides of DEFAULTS.  `nt` sets the
    step count through itmax = (nt + 0.5) * dt_fs unless texit_fs is given."""
    c = dict(DEFAULTS)
    unknown = set(kw) - set(c)
    if unknown:
        raise KeyError(f"unknown input keys: {sorted(unknown)}")
    c.update(kw)
    texit = c["texit_fs"]
    if texit is None:
        texit = (c["nt"] + 0.5) * dt_fs(c["q"], c["cfl"])
    box = c["box"]
    mbox = c["mbox"] or box
    eb = c["eps_bbox"]
    eb_s = f"({eb.real!r},{eb.imag!r})" if isinstance(eb, complex) else repr(float(eb))
    lines = [
        c["description"],
        "",
        '"Mesh options"',
        f"{c['i_pol']}\ti_pol",
        f"{c['q']!r}\tq (nm)",
        f"{c['ixmax']}\tixmax",
        f"{c['iymax']}\tiymax",
        f"{c['izmax']}\tizmax",
        f"{c['cfl']!r}\tCFL",
        f"{c['wzpml']!r}\twzPML",
        f"{c['npml']}\tnPML",
        f"{c['pmlzi']}\tpmlzi",
        f"{c['pmlzf']}\tpmlzf",
        "",
        '"Source options"',
        f"{c['plane_wave']}\tplane_wave",
        f"{c['tincr']!r} {c['tdecay']!r}\ttincr_PW,tdecay_PW",
        f"{c['epump']!r} {c['wlpump']!r} {c['pump_dt']!r} {c['def_pulse']}\tEpump,wlpump,pump_DT,def",
        f"{c['tfsf_o']}\ttfsf_o",
        f"{c['cosine']}\tcosine",
        "",
        '"Job options"',
        f"{c['no_downloads']} {c['output_time_step']}\tno_downloads,OUTPUT_TIME_STEP",
        f"{texit!r}\titmax (fs)",
        f"{c['hi']}\thi",
        f"{c['sdl'][0]} {c['sdl'][1]} {c['sdl'][2]}\tSDL_X,SDL_Y,SDL_Z",
        "",
        '"Dielectric constants"',
        f"{c['eps1']!r}\teps_cover",
        f"{eb_s}\teps_bbox",
        f"{c['eps3']!r}\teps_substrate",
        "",
        '"Fourier transformation options"',
        f"{box[0]} {mbox[0]}\tsxi mxi",
        f"{box[1]} {mbox[1]}\tsxf mxf",
        f"{box[2]} {mbox[2]}\tsyi myi",
        f"{box[3]} {mbox[3]}\tsyf myf",
        f"{box[4]} {mbox[4]}\tszi mzi",
        f"{box[5]} {mbox[5]}\tszf mzf",
        f"{c['nof']}\tnof",
        f"{c['dlambda']!r}\tdlambda",
        f"{c['lno']!r}\tlno",
        f"{c['to_fourier']} {c['sample_fourier']}\tto_Fourier,sample_Fourier",
        "",
    ]
    if c["precision"] is not None:
        lines += ['"Engine options"', f"{c['precision']}\tprecision (8: FP64 fields, 4: FP32 fields)", ""]
    if c["sdl_opts"] is not None:
        o = dict(SDL_DEFAULTS)
        unknown = set(c["sdl_opts"]) - set(o)
        if unknown:
            raise KeyError(f"unknown SDL keys: {sorted(unknown)}")
        o.update(c["sdl_opts"])
        lines += ['"Time evolution options"', f"{o['copy'][0]} {o['copy'][1]}\tSDL_copy_flag,SDL_copy_no"]
        for key in ("xi", "xf", "yi", "yf", "zi", "zf"):
            v = o[key]
            lines.append(f"{v[0]} {v[1]} {v[2]}\ts{key}")
        lines.append(f"{o['module_e']}\tE")
        for name, v in zip(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"), o["fields"]):
            lines.append(f"{v}\t{name}")
        lines.append("")
        if o["layers"] is not None:
            counts, dxs, dys, dzs = o["layers"]
            lines += ['"Extra parameters-SDLextended"', f"{counts[0]} {counts[1]} {counts[2]}"]
            lines += [str(d) for d in list(dxs) + list(dys) + list(dzs)]
            lines.append("")
    lines += [
        '"Slit Array"',
        f"{c['dh']}\tDh",
        f"{c['dx']}\tdx",
        "",
    ]
    return "\n".join(lines)
