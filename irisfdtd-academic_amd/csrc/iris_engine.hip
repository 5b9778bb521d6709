// iris_engine.hip -- C ABI (include/iris.h) of the MI355X FDTD engine.
//
// Replaces the time loop of IrisFDTD-Academic's driver()
// (reference code/driver.f90:338-500).  Device memory layout:
//   fields   E and H, one complex array per component, (nx+2)(ny+2)(nz+2)
//            cells with explicit ghost planes; x fastest with a padded pitch
//            so that ix = 1 starts on a 128-byte boundary; updated in place;
//   UPML     D (E side) and B (H side) per component for the bottom
//            [1,pmlzi] and top [pmlzf,nz] slabs, same xy pitch;
//   eps_inv  real (and, only if non-zero, imaginary) part on the structure
//            planes [hi,hf];
//   DFT      per box, the reference's emg(j)%fdata blocks back to back in
//            FP64 complex, plus a K-slot ring of face samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "../../include/iris.h"
#include "iris_adjoint.h"
#include "iris_incident.h"
#include "iris_kernels.h"
#include "iris_yee2p.h"

using namespace iris;

namespace {

// samples buffered before folding into emg: 32 with FP64 fields (each
// accumulator takes them one by one in the reference's order, so the ring
// length changes only how often it is stored and reloaded), 64 with FP32
// fields (the matrix-core fold)
constexpr int kDftSlots64 = 32, kDftSlots32 = 64;
// waves per workgroup of the matrix-core fold and the wavelength groups (16
// each) whose phases it stages at once (KMAX x 16 x 16 bytes a group)
constexpr int kFoldWaves = 16, kFoldGroups = 8;
template <typename T> constexpr int dft_slots() { return sizeof(T) == 8 ? kDftSlots64 : kDftSlots32; }

thread_local std::string g_create_error = "no error";

// Staging of iris_get_plane_noalloc: a fixed device array instead of the
// handle's growable buffer, for callers that may not allocate.  In host code
// `scratch` names a host-side stand-in; the device address of each GPU's
// copy comes from hipGetSymbolAddress, once per device (at iris_create, with
// that device current), and only that address is used.
constexpr int64_t kPlaneScratch = int64_t(1) << 20;  // complex values (16 MiB of double2)
static __device__ __align__(256) double2 scratch[kPlaneScratch];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Resolve the current device's address of `scratch` (no-op when known).
hipError_t scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[dev] = ptr;
  return e;
}

struct DeviceBox {
  BoxGeom g{};
  int active_mask = 0;
  int64_t npts_ref[3] = {1, 1, 1};  // what the reference allocates
  double2* emg = nullptr;
  void* samp = nullptr;
};

// single-step chunks of one length inside the rest list of a two-step pass
struct RestGroup {
  int len;          // planes per chunk (the launches' zchunk)
  int off;          // first entry in the rest list
  int nlean, ngen;  // lean entries, then general ones
};

// chunk lists of a two-step pass: two-step chunks (z0 | length | export
// flags) and the chunks left to single steps, lean ones first
struct PassLists {
  const int* chunks2 = nullptr;
  int n2 = 0;
  const int* chunks2p = nullptr;  // two-step chunks reaching into the UPML slabs
  int n2p = 0;
  const int* rest = nullptr;
  int nlean = 0, ngen = 0;
  const unsigned char* zmask = nullptr;  // 1 on the planes of two-step chunks
  const RestGroup* groups = nullptr;      // rest chunks by length (null: one group of zchunk planes)
  int ngroups = 0;
  int rz0 = 1, rz1 = 0;                   // record planes whose H of step n+1 the two-step pass exports
};

// device copy of the lists of one two-step classification
struct ListSet {
  bool on = false;
  int key = -1;
  int gen = -1;          // pass_gen of the layout the lists were made for
  int rz0 = 1, rz1 = 0;  // record planes inside two-step chunks (their H of step n+1 exported)
  int* chunks2 = nullptr;
  int n2 = 0;
  int* chunks2p = nullptr;  // two-step chunks of the UPML variant (record lists only)
  int n2p = 0;
  int* rest = nullptr;
  int nlean = 0, ngen = 0;
  unsigned char* zmask = nullptr;
  PassLists view() const {
    PassLists L;
    L.chunks2 = chunks2; L.n2 = n2; L.rest = rest; L.nlean = nlean; L.ngen = ngen; L.zmask = zmask;
    L.chunks2p = chunks2p; L.n2p = n2p;
    L.rz0 = rz0; L.rz1 = rz1;
    return L;
  }
};

}  // namespace

// Layout of the step passes: chunk lengths, which planes take the two-step
// pass and its UPML variant, the dealing of tiles to XCDs.  Every layout
// gives the same results bit for bit, except that with a Bloch phase the
// two-step pass (and its UPML variant) agree with single steps to rounding.
struct PassConfig {
  int zc = 0;         // planes per single-step z-chunk (0: enough chunks for ~4096 workgroups, at most 32 planes)
  int two = -1;       // two-step pass: 0 off, 1 wherever it applies, -1 where its tiles waste fewer lanes
  int upml = -1;      // its UPML variant: 0 off, 1 wherever it applies, -1 where its launch fills the CUs
  int zc2 = 0;        // planes per two-step chunk (0: 2*zc)
  int xcd_rows = 0;   // tile rows per XCD run of the single-step pass (0: identity order)
  int xcd_rows2 = -1; // the same for the two-step pass (-1: 2 with two cells a lane, else 4)
  int w2 = 0;         // two-step tile width: 64, 32 or 128 (two cells a lane, FP32 fields); 0: by the grid
  int w2p = 0;        // the same for its UPML variant (0: 32)
  int nt2 = -1;       // two-step pass streams its outputs past the L2: 0 no, 1 yes, -1 up to 160 x 160 cells a plane or with two cells a lane
  int p2ox = 0;       // two cells a lane: output cells a row of the main tiles (124, 60, 28; 0: by the grid)
};

struct iris_handle {
  iris_params p{};
  int prec = 8;
  size_t csize = 16;  // bytes per complex value of the fields
  hipStream_t stream = nullptr;
  std::string err = "no error";
  int device = 0;
  // layout
  int xoff = 0;
  int64_t sx = 0, sy = 0, sxy = 0, ncell = 0;
  // device buffers
  // fields: one allocation per field, components cs apart
  void* fe = nullptr;       // E (current)
  void* fh = nullptr;       // H (current)
  void* ge = nullptr;       // fused pass: the other E and H buffers
  void* gh = nullptr;
  int64_t cs = 0;
  // UPML aux: bottom slab then top slab, components ps0/ps1 apart
  void* d = nullptr;
  void* b = nullptr;
  void* dn = nullptr;
  void* bn = nullptr;
  int64_t ps0 = 0, ps1 = 0, pb1 = 0;
  bool fused = true;        // fused E+H pass over double-buffered fields
  double* diag_buf = nullptr; // partials and result of iris_diagnostics
  void* plane_buf = nullptr; // staging of iris_get_plane
  int64_t plane_cap = 0;      // its capacity in complex values
  int xcd_rows = 0;         // tile rows per XCD run of the fused pass (0: identity order)
  dim3 blk_fused, grd_fused;
  TfsfConst* tfc = nullptr; // device copy of the TF/SF constants
  int* chunks = nullptr;    // first plane of each fused z-chunk: bulk ones, then general ones
  int nbulk = 0, ngen = 0;
  // two steps per pass (k_yee2) over the chunks whose planes z0-1..z1+2 are
  // plain; the others take two single steps (lean ones, then general ones)
  bool two_step = false;
  bool bloch = false;       // Bloch phase in x or y (the two-step pass's image lanes carry it)
  void* ke = nullptr;       // third E and H buffers (step n+2 of a two-step pass)
  void* kh = nullptr;
  int* chunks2 = nullptr;   // two-step chunks (z0 | export flags)
  int n2 = 0;
  int64_t planes2 = 0;      // planes advanced by the two-step pass
  int* chunks2p = nullptr;  // two-step chunks reaching into the UPML slabs (k_yee2<..., PML>)
  int n2p = 0;
  int64_t planes2p = 0, upml2p = 0;  // their planes, and the UPML planes among them
  int ty2p = 0;             // tile rows of the UPML variant
  void* kd = nullptr;       // third D and B buffers (the UPML variant's output)
  void* kb = nullptr;
  int* chunks_rest = nullptr; // the other chunks: by length, lean ones then general ones
  int nrest_lean = 0, nrest_gen = 0;
  std::vector<RestGroup> rest_groups;
  unsigned char* zmask = nullptr; // 1 on the planes of two-step chunks
  int tx2 = 0, ty2 = 0;
  int xcd_rows2 = 4;        // tile rows per XCD run of the two-step pass
  int zc2 = 0;              // planes per plain two-step chunk
  int ncu = 256;            // compute units of the device
  int t2w = 64;             // two-step tile width (footprint cells per row)
  int t2wp = 64;            // the same for its UPML variant
  int tx2p = 0;             // tile columns of the UPML variant
  // two cells a lane (t2w = 128): main tiles (tx2 x ty2, p2ox cells a row
  // in rows of p2sl lanes, p2ns rows a wave) over x in 1..p2ox*tx2 and
  // narrow tiles (tx2n x ty2n, the same for p2oxn, p2sln, p2nsn) over the
  // columns from x0n + 1
  int p2ox = 0, p2sl = 0, p2ns = 0, p2oxn = 0, p2sln = 0, p2nsn = 0, tx2n = 0, ty2n = 0, x0n = 0;
  int nt2 = 0;              // two-step pass: streaming stores of its outputs
  int fold_persist = 1;     // FP32 fold: persistent workgroups when nof <= 128 (0: one workgroup per tile block)
  PassConfig pass;          // the layout the lists were built for
  int64_t pass_gen = 0;     // bumped by every rebuild of the pass lists
  // single steps of the two-step layout's other chunks when they come in
  // several lengths: one batched launch per class (lean, general) over the
  // length groups, with per-group arguments from a table per buffer
  // assignment (the buffers rotate between a few assignments)
  int* rest_wg[2] = {nullptr, nullptr};  // workgroup prefixes over the groups
  int rest_nwg[2] = {0, 0};
  struct RestTab {
    void* key[12];
    void* dev;
    std::vector<char> host;  // kept while the upload may still read it
  };
  std::vector<RestTab> rest_tabs;
  void* eps_re = nullptr;
  void* eps_im = nullptr;
  void* pml = nullptr;
  double* fw = nullptr;
  double2* ph = nullptr;
  double* flux = nullptr;
  DeviceBox box[2];
  int nsamp = 0;            // samples waiting in the ring
  int samp_n0 = 0;          // DFT phase index of the first waiting sample
  // steps as a captured HIP graph: one period of steps (the sample ring
  // fills once, the buffers come back to their assignment) captured at step
  // g_it0 and replayed at later steps with the TF/SF step shift set
  bool graph_on = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_it0 = 0;
  void* g_key[12] = {};     // buffer assignment the graph was captured with
  void* g_warm[12] = {};    // assignment whose period last ran eagerly (tables made)
  bool g_warmed = false;
  int64_t g_replays = 0;
  int it = 0;
  int zchunk = 1;
  dim3 blk_yee, grd_yee, blk_tf, grd_tf;
  double dtcq2 = 0;
  std::vector<void*> allocs;
  // profiling: event pairs around launches, drained in batches
  struct Pending { int k; hipEvent_t a, b; };
  bool prof = false;
  std::vector<Pending> pend;
  std::vector<hipEvent_t> pool;
  double prof_ms[IRIS_NKERNEL] = {};
  int64_t prof_n[IRIS_NKERNEL] = {};
  // gradient: forward snapshots every ck_seg steps (E, H, D, B back to back)
  int ck_seg = 0;
  std::vector<void*> ck;    // snapshot k holds the state after step k*ck_seg
  std::vector<char> ck_ok;
  void* ck_final = nullptr; // the state at nt while iris_adjoint recomputes
  // adjoint fields and work buffers (allocated by the first iris_adjoint)
  void* ae = nullptr;
  void* ah = nullptr;
  void* ad = nullptr;
  void* ab = nullptr;
  void* atmp = nullptr;
  void* ae2 = nullptr;      // fused adjoint step: the other adjoint buffers
  void* ah2 = nullptr;
  void* ad2 = nullptr;
  void* ab2 = nullptr;
  void* ae3 = nullptr;      // third adjoint E and H (two-step adjoint passes)
  void* ah3 = nullptr;
  // two-step lists of the gradient runs: the structure planes hi-1..hf and
  // the TF/SF planes stay in single-step chunks (their H is recorded), and
  // for the adjoint sweep also the planes of weighted faces (their sources
  // enter between the two steps)
  ListSet rl;               // forward steps that record H (face independent)
  ListSet al;               // adjoint steps (for one box and face mask)
  // H records of the last steps of the forward run: H of steps t0..nt-1,
  // kept while stepping, so that those steps are never recomputed
  int tail_t0 = 0;
  int tail_n = 0;
  void* tail_rec = nullptr;
  std::vector<char> tail_ok;
  bool adj_fused = true;    // fused adjoint step (IRIS_ADJ_FUSED=0: one kernel per part)
  int adj_ty = 0;           // tile rows of the general adjoint pass (0: 7 with FP32 fields, 8 with FP64)
  void* hrec = nullptr;     // H on planes hi-1..hf for each step of a segment
  int64_t rec_plane = 0;    // complex values per component of one record
  int rec_cap = 0;          // records hrec holds
  double2* adj_g = nullptr; // dJ/d(emg) of the objective's box
  int64_t adj_g_cap = 0;
  double2* grad = nullptr;  // accumulated gradient on the structure cells
  double* wdev = nullptr;   // objective weights (6*nof)
  std::vector<void*> adj_allocs;
};

// A batch of handles stepped together by the batched fused pass.
struct iris_batch {
  std::vector<iris_handle*> hs;     // all handles of the batch
  std::vector<int64_t> pass_gen;    // their pass_gen when the batch was made (its tables hold those lists)
  // the handles' own layouts, put back by iris_batch_destroy (the batch lays
  // them out for itself: see batch_layout)
  std::vector<PassConfig> saved;
  std::vector<iris_handle*> one;    // those the batched single-step pass advances
  std::vector<iris_handle*> own;    // those that take their own two-step passes
  int prec = 4;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::string err = "no error";
  int zc = 32;                      // planes per z-chunk of the batched passes
  std::vector<void*> allocs;        // device memory of the batch
  std::vector<int*> lean, gen;      // per instance: chunk lists
  std::vector<int> nlean, ngen;
  int* wg[2] = {nullptr, nullptr};  // per class (lean, general): workgroup prefix over the instances
  int nwg[2] = {0, 0};
  void* inst = nullptr;             // BatchInst<T>[n]
  void* ent = nullptr;              // BatchSamp<T>[nent]
  int64_t* ent_start = nullptr;
  int nent = 0;
  int64_t npts = 0;
  std::vector<char> host_inst, host_ent;
  // the two-step handles (`own`), batched: pairs of steps by one launch per
  // kernel over all of them.  Tables per pair parity (the pair runs from the
  // current buffers to the third ones, the next pair back).
  // by kernels: Bloch phase (bit 0), 32-cell tiles of the plain pass (bit 1) and of its UPML variant (bit 2),
  // 128-cell tiles of two cells a lane of the plain pass (bit 3)
  std::vector<iris_handle*> own2[16];
  // table entries of a group's launch: (handle, part), the part being the
  // full (0) or narrow (1) tiles of the two-cells-a-lane kernel, else 0
  std::vector<std::pair<iris_handle*, int>> y2e[16][2];
  void* y2[16][2] = {};               // Batch2Inst<T>[2][entries]: [group][upml variant], pair parity outer
  int* y2wg[16][2] = {};              // workgroup prefixes
  int y2nwg[16][2] = {};
  void* rinst = nullptr;              // BatchInst<T>[2][nr]: their single-step chunks, per (handle, length group)
  int nr = 0;
  int* rwg[2] = {nullptr, nullptr};   // per class (lean, general)
  int rnwg[2] = {0, 0};
  void* ent2 = nullptr;               // BatchSamp<T>[2][nent2]: step n+1 (masked), step n+2
  int64_t* ent2_start = nullptr;
  int nent2 = 0;
  int64_t npts2 = 0;
  // the pairs of one period of steps as a captured HIP graph (as a handle
  // alone does): captured at step g_it0 at pair parity 0 and replayed there
  // with every two-step handle's TF/SF step shift set.  The launches read
  // the batch's tables, which every iris_batch_step rewrites in place for
  // the handles' buffers at that call, when the pair parity starts at 0.
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int g_it0 = 0;
  int64_t g_replays = 0;
};

namespace {

int fail(iris_handle* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf; else g_create_error = buf;
  return code;
}

#define HIP_TRY(h, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return fail((h), IRIS_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

int dalloc(iris_handle* h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(h, IRIS_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
  }
  h->allocs.push_back(*p);
  e = hipMemsetAsync(*p, 0, bytes, h->stream);
  if (e != hipSuccess) return fail(h, IRIS_EDEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
  return 0;
}

int pow2ceil(int v) { int p = 1; while (p < v) p <<= 1; return p; }

// ---- profiling -------------------------------------------------------------
hipEvent_t take_event(iris_handle* h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void prof_drain(iris_handle* h) {
  for (auto& p : h->pend) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      h->prof_ms[p.k] += ms;
      h->prof_n[p.k] += 1;
    }
    h->pool.push_back(p.a);
    h->pool.push_back(p.b);
  }
  h->pend.clear();
}

// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
  iris_handle* h;
  int k;
  hipEvent_t a = nullptr;
  ProfScope(iris_handle* h_, int k_) : h(h_), k(k_) {
    if (h->prof && (a = take_event(h))) (void)hipEventRecord(a, h->stream);
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = take_event(h);
    if (!b) { h->pool.push_back(a); return; }
    (void)hipEventRecord(b, h->stream);
    h->pend.push_back({k, a, b});
    if (h->pend.size() >= 1024) prof_drain(h);
  }
};

// Bloch factor exp(+-i*ak) as the reference evaluates cdexp(+-ci*ak)
void bloch(const double ak[2], int sign, double out[2]) {
  // ci*ak = (0*a - 1*b, 0*b + 1*a)
  double re = 0.0 * ak[0] - 1.0 * ak[1], im = 0.0 * ak[1] + 1.0 * ak[0];
  if (sign < 0) { re = -re; im = -im; }
  double m = std::exp(re);
  out[0] = m * std::cos(im);
  out[1] = m * std::sin(im);
}

template <typename T>
StepArgs<T> step_args(const iris_handle* h) {
  using C = cplx<T>;
  StepArgs<T> a{};
  a.e = (C*)h->fe;
  a.h = (C*)h->fh;
  a.en = (C*)(h->fused ? h->ge : h->fe);
  a.hn = (C*)(h->fused ? h->gh : h->fh);
  a.d = (C*)h->d;
  a.b = (C*)h->b;
  a.dn = (C*)(h->fused ? h->dn : h->d);
  a.bn = (C*)(h->fused ? h->bn : h->b);
  a.cs = h->cs;
  a.ps0 = h->ps0;
  a.ps1 = h->ps1;
  a.pb1 = h->pb1;
  a.eps_re = (const T*)h->eps_re;
  a.eps_im = (const T*)h->eps_im;
  a.pml = (const PmlPlane<T>*)h->pml;
  a.nx = h->p.nx; a.ny = h->p.ny; a.nz = h->p.nz;
  a.sx = h->sx; a.sxy = h->sxy; a.xoff = h->xoff;
  a.pmlzi = h->p.pmlzi; a.pmlzf = h->p.pmlzf; a.hi = h->p.hi; a.hf = h->p.hf;
  for (int i = 0; i < 3; ++i) { a.a1[i] = (T)h->p.epsa_inv[0][i][i]; a.a3[i] = (T)h->p.epsa_inv[1][i][i]; }
  a.e02 = (T)(h->p.eps0 * h->p.eps0);
  a.dtcq2 = (T)h->dtcq2;
  double f[2];
  bloch(h->p.ak[0], +1, f); a.fxp = mk<T>((T)f[0], (T)f[1]);
  bloch(h->p.ak[1], +1, f); a.fyp = mk<T>((T)f[0], (T)f[1]);
  bloch(h->p.ak[2], +1, f); a.fzp = mk<T>((T)f[0], (T)f[1]);
  bloch(h->p.ak[0], -1, f); a.fxm = mk<T>((T)f[0], (T)f[1]);
  bloch(h->p.ak[1], -1, f); a.fym = mk<T>((T)f[0], (T)f[1]);
  bloch(h->p.ak[2], -1, f); a.fzm = mk<T>((T)f[0], (T)f[1]);
  a.zchunk = h->zchunk;
  a.nt_out = h->nt2;
  return a;
}

// TF/SF constants as the reference evaluates them (code/tfsf.f90:262-290)
TfsfConst tfsf_const(const iris_handle* h) {
  const iris_params& p = h->p;
  TfsfConst t{};
  t.q = p.q; t.ux = p.ux; t.uz = p.uz; t.zopump = p.zopump;
  t.ko = (2.0 * p.pi) / (p.wlpump * p.f_abohr);
  t.dk = p.gpump * p.q;
  double eps_ms;
  if (p.iz_min > p.hi) { eps_ms = p.eps[2]; t.sign_kz = -1.0; } else { eps_ms = p.eps[0]; t.sign_kz = 1.0; }
  t.sq = std::sqrt(eps_ms);
  t.eps_ms = eps_ms;
  t.c0dt = p.c0 * p.dt;
  t.amp = p.eamp * p.q;
  t.hamp = t.amp * ((p.dt * p.c0) / p.q);
  t.win = p.win;
  t.ein = p.eps_in[0];
  t.dtcq2 = h->dtcq2 * 1.0;
  t.ixo = p.ixo;
  t.iz_min = p.iz_min;
  t.i_pol = p.i_pol;
  t.cosine = p.cosine;
  t.mode2d = (p.nx == 1 || p.ny == 1) ? 1 : 0;
  t.on = p.em_source ? 1 : 0;
  return t;
}

// StepArgs with explicit buffers (the two-step pass chains passes through
// the three field buffers)
template <typename T>
StepArgs<T> step_args_on(const iris_handle* h, void* e, void* hf, void* en, void* hn, void* d, void* b, void* dn,
                         void* bn) {
  using C = cplx<T>;
  StepArgs<T> a = step_args<T>(h);
  a.e = (C*)e; a.h = (C*)hf; a.en = (C*)en; a.hn = (C*)hn;
  a.d = (C*)d; a.b = (C*)b; a.dn = (C*)dn; a.bn = (C*)bn;
  return a;
}

// single-step fused launches over lists of lean and general chunks
template <typename T>
void launch_fused_lists(iris_handle* h, const StepArgs<T>& a, int it, const int* lean, int nlean, const int* gen,
                        int ngen) {
  TileMap tm{(int)h->grd_fused.x, (int)h->grd_fused.y, 0, (int)h->grd_fused.x * h->xcd_rows};
  if (nlean) {
    tm.nblk = tm.tx * tm.ty * nlean;
    hipLaunchKernelGGL((k_yee_fused<T, false>), dim3(tm.nblk), h->blk_fused, 0, h->stream, a,
                       (const TfsfConst*)h->tfc, it, lean, tm);
  }
  if (ngen) {
    tm.nblk = tm.tx * tm.ty * ngen;
    hipLaunchKernelGGL((k_yee_fused<T, true>), dim3(tm.nblk), h->blk_fused, 0, h->stream, a,
                       (const TfsfConst*)h->tfc, it, gen, tm);
  }
}

template <typename T>
int dft_sample(iris_handle* h, int n, const unsigned char* xmask = nullptr);
template <typename T>
int dft_flush(iris_handle* h);

// One launch of the two-step kernel (variant PH, PML) at the handle's tile
// width over nch chunks (tile map tm; with two cells a lane the handle's
// full and narrow tiles instead).
template <typename T, bool PH, bool PML>
void launch_yee2(const iris_handle* h, const StepArgs<T>& a, cplx<T>* eb, cplx<T>* hb, cplx<T>* db,
                 const int* chunks, const TileMap& tm, int nch, const Yee2Samples<T>& smp, int slot, int slot2) {
  if constexpr (sizeof(T) == 4 && !PML) {
    if (h->t2w == 128) {
      const P2Part p0{TileMap{h->tx2, h->ty2, h->tx2 * h->ty2 * nch, h->tx2 * h->xcd_rows2, 0}, h->p2ox, h->p2sl,
                      h->p2ns};
      const P2Part p1{TileMap{h->tx2n, h->ty2n, h->tx2n * h->ty2n * nch, h->tx2n * h->xcd_rows2, h->x0n}, h->p2oxn,
                      h->p2sln, h->p2nsn};
      const int npad0 = p1.tm.nblk ? (p0.tm.nblk + kNumXcd - 1) / kNumXcd * kNumXcd : p0.tm.nblk;
      const unsigned nblk = (unsigned)(npad0 + p1.tm.nblk);
      if (nblk)
        hipLaunchKernelGGL((k_yee2p<PH>), dim3(nblk), dim3(64, kP2R), 0, h->stream, a, (float2*)eb, (float2*)hb,
                           chunks, p0, p1, npad0, smp, slot, slot2);
      return;
    }
  }
  (void)nch;
  const dim3 grid((unsigned)tm.nblk);
  if ((PML ? h->t2wp : h->t2w) == 32)
    hipLaunchKernelGGL((k_yee2<T, PH, PML, 32>), grid, dim3(kT2X, Yee2Shape<T, PML, 32>::R), 0, h->stream, a, eb,
                       hb, db, chunks, tm, smp, slot, slot2);
  else
    hipLaunchKernelGGL((k_yee2<T, PH, PML, 64>), grid, dim3(kT2X, Yee2Shape<T, PML, 64>::R), 0, h->stream, a, eb,
                       hb, db, chunks, tm, smp, slot, slot2);
}

// Argument table of the handle's single-step length groups for its current
// buffers (entry g: a[0] = step n+1 from the current into the second
// buffers, a[1] = step n+2 from the second into the third), made on first
// use of a buffer assignment.
template <typename T>
int rest_table(iris_handle* h, const BatchInst<T>** out) {
  void* key[12] = {h->fe, h->fh, h->ge, h->gh, h->ke, h->kh, h->d, h->b, h->dn, h->bn, h->kd, h->kb};
  for (const auto& t : h->rest_tabs)
    if (std::equal(key, key + 12, t.key)) {
      *out = (const BatchInst<T>*)t.dev;
      return 0;
    }
  const int n = (int)h->rest_groups.size();
  iris_handle::RestTab t;
  std::copy(key, key + 12, t.key);
  t.host.assign(sizeof(BatchInst<T>) * n, 0);
  BatchInst<T>* I = (BatchInst<T>*)t.host.data();
  const StepArgs<T> a0 = step_args_on<T>(h, h->fe, h->fh, h->ge, h->gh, h->d, h->b, h->dn, h->bn);
  const StepArgs<T> a1 = step_args_on<T>(h, h->ge, h->gh, h->ke, h->kh, h->dn, h->bn, h->kd, h->kb);
  const int tx = (int)h->grd_fused.x, ty = (int)h->grd_fused.y;
  for (int g = 0; g < n; ++g) {
    const RestGroup& G = h->rest_groups[g];
    I[g].a[0] = a0;
    I[g].a[1] = a1;
    I[g].a[0].zchunk = I[g].a[1].zchunk = G.len;
    I[g].tp = (const TfsfConst*)h->tfc;
    I[g].chunks[0] = h->chunks_rest + G.off;
    I[g].chunks[1] = h->chunks_rest + G.off + G.nlean;
    I[g].tm[0] = TileMap{tx, ty, tx * ty * G.nlean, tx * h->xcd_rows};
    I[g].tm[1] = TileMap{tx, ty, tx * ty * G.ngen, tx * h->xcd_rows};
  }
  int rc = dalloc(h, &t.dev, t.host.size());
  if (rc) return rc;
  HIP_TRY(h, hipMemcpyAsync(t.dev, t.host.data(), t.host.size(), hipMemcpyHostToDevice, h->stream));
  h->rest_tabs.push_back(std::move(t));
  *out = (const BatchInst<T>*)h->rest_tabs.back().dev;
  return 0;
}

// Steps it and it+1 in one pass: the two-step kernel over its chunks (step n
// from the current buffers, step n+2 into the third), the other chunks by two
// single-step launches (current -> second buffer -> third), then the DFT
// samples of both steps.
template <typename T>
int launch_double(iris_handle* h, int it, const PassLists& L, bool dft) {
  using C = cplx<T>;
  const iris_params& p = h->p;
  const bool s1 = dft && it >= p.to_fourier && (it % p.sample_fourier) == 0;
  const bool s2 = dft && it + 1 >= p.to_fourier && ((it + 1) % p.sample_fourier) == 0;
  // ring slots: step n+1 (every face of the two-step kernels' own cells),
  // step n+2 (their x faces; not when step n+1 fills the ring, whose fold
  // must come first)
  const int slot = s1 ? h->nsamp : -1;
  const int slot2 = (s2 && h->nsamp + (s1 ? 1 : 0) < dft_slots<T>()) ? h->nsamp + (s1 ? 1 : 0) : -1;
  Yee2Samples<T> smp{};
  smp.nbox = p.nbox;
  for (int b = 0; b < p.nbox; ++b) {
    smp.g[b] = h->box[b].g;
    smp.samp[b] = h->box[b].g.ntot ? (C*)h->box[b].samp : nullptr;
  }
  const RestGroup one{h->zchunk, 0, L.nlean, L.ngen};
  const RestGroup* groups = L.groups ? L.groups : &one;
  const int ngroups = L.groups ? L.ngroups : 1;
  // the handle's own lists in several lengths: one launch per class over
  // all groups (k = 0: step n+1 into the second buffers, 1: step n+2)
  const BatchInst<T>* tab = nullptr;
  if (L.groups == h->rest_groups.data() && ngroups > 1) {
    int rc = rest_table<T>(h, &tab);
    if (rc) return rc;
  }
  auto rest = [&](StepArgs<T> a, int step) {
    if (tab) {
      const int k = step == it ? 0 : 1;
      if (h->rest_nwg[0])
        hipLaunchKernelGGL((k_yee_fused_batch<T, false>), dim3(h->rest_nwg[0]), h->blk_fused, 0, h->stream, tab,
                           (const int*)h->rest_wg[0], ngroups, k, step);
      if (h->rest_nwg[1])
        hipLaunchKernelGGL((k_yee_fused_batch<T, true>), dim3(h->rest_nwg[1]), h->blk_fused, 0, h->stream, tab,
                           (const int*)h->rest_wg[1], ngroups, k, step);
      return;
    }
    for (int g = 0; g < ngroups; ++g) {
      const RestGroup& G = groups[g];
      if (G.nlean + G.ngen == 0) continue;
      a.zchunk = G.len;
      launch_fused_lists<T>(h, a, step, L.rest + G.off, G.nlean, L.rest + G.off + G.nlean, G.ngen);
    }
  };
  if (L.nlean + L.ngen) {
    ProfScope ps(h, IRIS_K_YEE_REST);
    rest(step_args_on<T>(h, h->fe, h->fh, h->ge, h->gh, h->d, h->b, h->dn, h->bn), it);
  }
  if (L.n2) {
    ProfScope ps(h, IRIS_K_YEE2);
    StepArgs<T> a2 = step_args_on<T>(h, h->fe, h->fh, h->ke, h->kh, h->d, h->b, h->dn, h->bn);
    a2.rec_z0 = L.rz0;
    a2.rec_z1 = L.rz1;
    const TileMap tm{h->tx2, h->ty2, h->tx2 * h->ty2 * L.n2, h->tx2 * h->xcd_rows2};
    if (h->bloch)
      launch_yee2<T, true, false>(h, a2, (C*)h->ge, (C*)h->gh, nullptr, L.chunks2, tm, L.n2, smp, slot, slot2);
    else
      launch_yee2<T, false, false>(h, a2, (C*)h->ge, (C*)h->gh, nullptr, L.chunks2, tm, L.n2, smp, slot, slot2);
  }
  if (L.n2p) {
    // the UPML variant: D and B of step n+2 into the third aux buffers
    ProfScope ps(h, IRIS_K_YEE2_PML);
    const StepArgs<T> a2 = step_args_on<T>(h, h->fe, h->fh, h->ke, h->kh, h->d, h->b, h->kd, h->kb);
    const TileMap tm{h->tx2p, h->ty2p, h->tx2p * h->ty2p * L.n2p, h->tx2p * h->xcd_rows2};
    if (h->bloch)
      launch_yee2<T, true, true>(h, a2, (C*)h->ge, (C*)h->gh, (C*)h->dn, L.chunks2p, tm, L.n2p, smp, slot, slot2);
    else
      launch_yee2<T, false, true>(h, a2, (C*)h->ge, (C*)h->gh, (C*)h->dn, L.chunks2p, tm, L.n2p, smp, slot, slot2);
  }
  if (L.nlean + L.ngen) {
    ProfScope ps(h, IRIS_K_YEE_REST);
    rest(step_args_on<T>(h, h->ge, h->gh, h->ke, h->kh, h->dn, h->bn, h->kd, h->kb), it + 1);
  }
  HIP_TRY(h, hipGetLastError());
  if (s1) {
    // step it outside the two-step chunks is in the second buffer
    const StepArgs<T> ab = step_args_on<T>(h, h->ge, h->gh, h->ge, h->gh, h->d, h->b, h->d, h->b);
    if (h->nsamp == 0) h->samp_n0 = it + 1 - p.to_fourier;
    for (int b = 0; b < p.nbox; ++b) {
      DeviceBox& B = h->box[b];
      if (B.g.ntot == 0) continue;
      ProfScope ps(h, IRIS_K_DFT_SAMPLE);
      hipLaunchKernelGGL(k_dft_sample_masked<T>, dim3((unsigned)((B.g.ntot + 255) / 256)), dim3(256), 0, h->stream,
                         ab, B.g, (C*)B.samp, h->nsamp, L.zmask);
    }
    HIP_TRY(h, hipGetLastError());
    h->nsamp++;
    if (h->nsamp == dft_slots<T>()) {
      int rc = dft_flush<T>(h);
      if (rc) return rc;
    }
  }
  std::swap(h->fe, h->ke);
  std::swap(h->fh, h->kh);
  std::swap(h->d, h->kd);
  std::swap(h->b, h->kb);
  h->it = it;
  if (s2) {
    // the x faces of the two-step planes are in slot2 already
    int rc = dft_sample<T>(h, it + 2 - p.to_fourier, slot2 >= 0 ? L.zmask : nullptr);
    if (rc) return rc;
  }
  h->it = it + 1;
  return 0;
}

template <typename T>
int launch_step(iris_handle* h, int it) {
  StepArgs<T> a = step_args<T>(h);
  if (h->fused) {
    {
      ProfScope ps(h, IRIS_K_YEE);
      TileMap tm{(int)h->grd_fused.x, (int)h->grd_fused.y, 0, (int)h->grd_fused.x * h->xcd_rows};
      if (h->nbulk) {
        tm.nblk = tm.tx * tm.ty * h->nbulk;
        hipLaunchKernelGGL((k_yee_fused<T, false>), dim3(tm.nblk), h->blk_fused, 0, h->stream, a,
                           (const TfsfConst*)h->tfc, it, (const int*)h->chunks, tm);
      }
      if (h->ngen) {
        tm.nblk = tm.tx * tm.ty * h->ngen;
        hipLaunchKernelGGL((k_yee_fused<T, true>), dim3(tm.nblk), h->blk_fused, 0, h->stream, a,
                           (const TfsfConst*)h->tfc, it, (const int*)h->chunks + h->nbulk, tm);
      }
    }
    HIP_TRY(h, hipGetLastError());
    std::swap(h->fe, h->ge);
    std::swap(h->fh, h->gh);
    std::swap(h->d, h->dn);
    std::swap(h->b, h->bn);
    return 0;
  }
  {
    ProfScope ps(h, IRIS_K_EPASS);
    hipLaunchKernelGGL(k_epass<T>, h->grd_yee, h->blk_yee, 0, h->stream, a);
  }
  {
    ProfScope ps(h, IRIS_K_HPASS);
    hipLaunchKernelGGL(k_hpass<T>, h->grd_yee, h->blk_yee, 0, h->stream, a);
  }
  if (h->p.em_source) {
    ProfScope ps(h, IRIS_K_TFSF);
    hipLaunchKernelGGL(k_tfsf<T>, h->grd_tf, h->blk_tf, 0, h->stream, a, (const TfsfConst*)h->tfc, it);
  }
  HIP_TRY(h, hipGetLastError());
  return 0;
}

template <typename T>
int dft_flush(iris_handle* h) {
  if (h->nsamp == 0) return 0;
  const int nof = h->p.nof;
  const int nk = h->nsamp;
  constexpr bool exact = sizeof(T) == 8;
  const int nrows = dft_slots<T>();
  hipLaunchKernelGGL(k_dft_phase, dim3((nof * nrows + 255) / 256), dim3(256), 0, h->stream, h->fw, nof, nk, nrows,
                     h->samp_n0, h->p.sample_fourier, exact ? 1.0 : h->p.dt, h->ph, (const TfsfConst*)h->tfc);
  for (int b = 0; b < h->p.nbox; ++b) {
    DeviceBox& B = h->box[b];
    if (B.g.ntot == 0) continue;
    ProfScope ps(h, IRIS_K_DFT_ACCUM);
    if constexpr (exact) {
      const dim3 grid((unsigned)((B.g.ntot + 255) / 256), 1);
      hipLaunchKernelGGL((k_dft_fold<cplx<T>, dft_slots<T>()>), grid, dim3(256), 0, h->stream, B.g,
                         (const cplx<T>*)B.samp, nk, (const double2*)h->ph, nof, h->p.dt, B.emg);
    } else {
      // 16 point-components a wave, kFoldWaves waves a workgroup (they share
      // each staged group of phases)
      int64_t tiles = 0;
      for (int q = 0; q < 3; ++q) tiles += (B.g.npts[q] + 15) / 16;
      const int64_t nwg = (tiles + kFoldWaves - 1) / kFoldWaves;
      if (h->fold_persist && (nof + 15) / 16 <= kFoldGroups) {
        // all wavelength groups staged at once: one persistent workgroup a
        // CU (its 128 KB of LDS allow no second)
        const unsigned grid = (unsigned)std::min<int64_t>(nwg, (int64_t)h->ncu);
        hipLaunchKernelGGL((k_dft_fold_mfma_p<dft_slots<T>(), kFoldWaves, kFoldGroups>), dim3(grid),
                           dim3(64 * kFoldWaves), 0, h->stream, B.g, (const float2*)B.samp, nk,
                           (const double2*)h->ph, nof, B.emg, tiles);
      } else {
        hipLaunchKernelGGL((k_dft_fold_mfma<dft_slots<T>(), kFoldWaves, kFoldGroups>), dim3((unsigned)nwg),
                           dim3(64 * kFoldWaves), 0, h->stream, B.g, (const float2*)B.samp, nk,
                           (const double2*)h->ph, nof, B.emg);
      }
    }
  }
  HIP_TRY(h, hipGetLastError());
  h->nsamp = 0;
  return 0;
}

template <typename T>
int dft_sample(iris_handle* h, int n, const unsigned char* xmask) {
  if (h->nsamp == 0) h->samp_n0 = n;
  StepArgs<T> a = step_args<T>(h);
  for (int b = 0; b < h->p.nbox; ++b) {
    DeviceBox& B = h->box[b];
    if (B.g.ntot == 0) continue;
    ProfScope ps(h, IRIS_K_DFT_SAMPLE);
    hipLaunchKernelGGL(k_dft_sample<T>, dim3((unsigned)((B.g.ntot + 255) / 256)), dim3(256), 0, h->stream, a, B.g,
                       (cplx<T>*)B.samp, h->nsamp, xmask);
  }
  HIP_TRY(h, hipGetLastError());
  h->nsamp++;
  if (h->nsamp == dft_slots<T>()) return dft_flush<T>(h);
  return 0;
}

// ---- steps as a captured HIP graph -----------------------------------------
// The launches of one period of P steps, P = ring slots x sample cadence (the
// sample ring fills and folds exactly once, and the rotating field buffers
// come back to the assignment they started from), are captured once and
// replayed.  Step indices reach the kernels as launch arguments; a replay
// s steps after the capture runs with the handle's TF/SF step shift set to
// s, which the TF/SF source and the DFT phases add to their step index.  The
// first period at a buffer assignment runs eagerly: it makes the per-
// assignment argument tables (an upload a capture cannot hold).

void graph_release(iris_handle* h) {
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->g_warmed = false;
}

void buffer_key(const iris_handle* h, void* key[12]) {
  void* k[12] = {h->fe, h->fh, h->ge, h->gh, h->ke, h->kh, h->d, h->b, h->dn, h->bn, h->kd, h->kb};
  std::copy(k, k + 12, key);
}

template <typename T> int graph_period(const iris_handle* h) { return dft_slots<T>() * h->p.sample_fourier; }

// Whether the P steps from `it` may go through the graph: a whole period
// fits before the call's last step (and, with two-step pairs, a single step
// still follows), the ring is empty, every step of it samples on cadence,
// and nothing records per-step timing or snapshots.
template <typename T>
bool graph_fits(const iris_handle* h, int it, int it_last) {
  const int P = graph_period<T>(h);
  return h->graph_on && !h->prof && h->ck_seg == 0 && h->nsamp == 0 && it >= h->p.to_fourier &&
         it_last - it + 1 >= P + (h->two_step ? 1 : 0);
}

// P steps from `it` by the eager launches (under capture: their recording)
template <typename T>
int period_steps(iris_handle* h, int it, const PassLists& L) {
  const int P = graph_period<T>(h);
  int rc;
  if (h->two_step) {
    for (int k = 0; k < P; k += 2)
      if ((rc = launch_double<T>(h, it + k, L, true))) return rc;
    return 0;
  }
  for (int k = 0; k < P; ++k) {
    const int s = it + k;
    if ((rc = launch_step<T>(h, s))) return rc;
    if (s >= h->p.to_fourier && (s % h->p.sample_fourier) == 0 &&
        (rc = dft_sample<T>(h, s + 1 - h->p.to_fourier)))
      return rc;
    h->it = s;
  }
  return 0;
}

// Steps it .. it+P-1: a replay of the captured period, or its capture and
// first launch, or an eager period that prepares the capture.  *shifted is
// set when the step shift was left non-zero.
template <typename T>
int graph_period_run(iris_handle* h, int it, const PassLists& L, bool* shifted) {
  const int P = graph_period<T>(h);
  void* key[12];
  buffer_key(h, key);
  if (h->gexec && std::equal(key, key + 12, h->g_key) && (it - h->g_it0) % h->p.sample_fourier == 0) {
    hipLaunchKernelGGL(k_set_shift, dim3(1), dim3(64), 0, h->stream, h->tfc, it - h->g_it0);
    HIP_TRY(h, hipGetLastError());
    *shifted = true;  // the caller resets the shift before any eager step, also after an error here
    HIP_TRY(h, hipGraphLaunch(h->gexec, h->stream));
    h->it = it + P - 1;
    h->g_replays++;
    return 0;
  }
  if (!(h->g_warmed && std::equal(key, key + 12, h->g_warm))) {
    std::copy(key, key + 12, h->g_warm);
    h->g_warmed = true;
    return period_steps<T>(h, it, L);
  }
  // capture; the host state advances as the steps are recorded, so it is
  // restored if the capture fails (nothing has run).  A replay of the old
  // graph may still be running: wait for it before its exec is destroyed.
  if (h->gexec) HIP_TRY(h, hipStreamSynchronize(h->stream));
  graph_release(h);
  if (*shifted) {
    hipLaunchKernelGGL(k_set_shift, dim3(1), dim3(64), 0, h->stream, h->tfc, 0);
    *shifted = false;
  }
  const int nsamp = h->nsamp, n0 = h->samp_n0, it_prev = h->it;
  HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = period_steps<T>(h, it, L);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  hipGraphExec_t x = nullptr;
  hipError_t ei = hipSuccess;
  if (!rc && e == hipSuccess) ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (rc || e != hipSuccess || ei != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    h->fe = key[0]; h->fh = key[1]; h->ge = key[2]; h->gh = key[3]; h->ke = key[4]; h->kh = key[5];
    h->d = key[6]; h->b = key[7]; h->dn = key[8]; h->bn = key[9]; h->kd = key[10]; h->kb = key[11];
    h->nsamp = nsamp;
    h->samp_n0 = n0;
    h->it = it_prev;
    if (rc) return rc;
    return fail(h, IRIS_EDEVICE, "graph capture of steps %d..%d: %s", it, it + P - 1,
                hipGetErrorString(e != hipSuccess ? e : ei));
  }
  h->graph = g;
  h->gexec = x;
  h->g_it0 = it;
  std::copy(key, key + 12, h->g_key);
  HIP_TRY(h, hipGraphLaunch(h->gexec, h->stream));
  return 0;
}

template <typename T>
int run_steps(iris_handle* h, int it_first, int it_last) {
  int it = it_first;
  bool shifted = false;  // the TF/SF step shift is non-zero (reset before eager steps)
  auto unshift = [&]() -> int {
    if (!shifted) return 0;
    hipLaunchKernelGGL(k_set_shift, dim3(1), dim3(64), 0, h->stream, h->tfc, 0);
    shifted = false;
    HIP_TRY(h, hipGetLastError());
    return 0;
  };
  // two steps per pass while at least three steps remain: every call ends
  // with a single-step pass, so the previous step's E and H stay whole
  PassLists L;
  L.chunks2 = h->chunks2;
  L.n2 = h->n2;
  L.rest = h->chunks_rest;
  L.nlean = h->nrest_lean;
  L.ngen = h->nrest_gen;
  L.zmask = h->zmask;
  L.chunks2p = h->chunks2p;
  L.n2p = h->n2p;
  L.groups = h->rest_groups.data();
  L.ngroups = (int)h->rest_groups.size();
  while (graph_fits<T>(h, it, it_last)) {
    int rc = graph_period_run<T>(h, it, L, &shifted);
    if (rc) {
      (void)unshift();
      return rc;
    }
    it += graph_period<T>(h);
  }
  if (int rc = unshift()) return rc;
  while (h->two_step && it_last - it + 1 >= 3) {
    if (graph_fits<T>(h, it, it_last)) {
      // the ring has just emptied: the rest of the call in periods
      int rc = run_steps<T>(h, it, it_last);
      return rc;
    }
    int rc = launch_double<T>(h, it, L, true);
    if (rc) return rc;
    it += 2;
  }
  for (; it <= it_last; ++it) {
    if (graph_fits<T>(h, it, it_last)) return run_steps<T>(h, it, it_last);
    int rc = launch_step<T>(h, it);
    if (rc) return rc;
    if (it >= h->p.to_fourier && (it % h->p.sample_fourier) == 0) {
      rc = dft_sample<T>(h, it + 1 - h->p.to_fourier);
      if (rc) return rc;
    }
    h->it = it;
  }
  return 0;
}

int flush(iris_handle* h) { return h->prec == 8 ? dft_flush<double>(h) : dft_flush<float>(h); }

int check_params(const iris_params* p) {
  if (!p) return fail(nullptr, IRIS_EINVAL, "params is NULL");
  if (p->abi_version != IRIS_ABI_VERSION)
    return fail(nullptr, IRIS_EINVAL, "abi_version %d, library expects %d", p->abi_version, IRIS_ABI_VERSION);
  if (p->precision != 4 && p->precision != 8) return fail(nullptr, IRIS_EINVAL, "precision must be 4 or 8");
  if (p->nx < 1 || p->ny < 1 || p->nz < 3) return fail(nullptr, IRIS_EINVAL, "grid %d x %d x %d too small", p->nx, p->ny, p->nz);
  if ((int64_t)p->nx * p->ny * p->nz > (int64_t)1 << 33) return fail(nullptr, IRIS_EINVAL, "grid too large");
  if (p->pmlzi < 2 || p->pmlzf <= p->pmlzi || p->pmlzf >= p->nz)
    return fail(nullptr, IRIS_EINVAL, "need 2 <= pmlzi < pmlzf < nz (got %d, %d, nz=%d)", p->pmlzi, p->pmlzf, p->nz);
  if (p->hi < 1 || p->hf < p->hi || p->hf > p->nz) return fail(nullptr, IRIS_EINVAL, "structure [%d,%d] outside [1,nz]", p->hi, p->hf);
  if (p->em_source) {
    if (p->plane_wave != 1) return fail(nullptr, IRIS_EUNSUP, "only the Gaussian TF/SF source (plane_wave=1) is supported");
    if (p->i_pol != 1 && p->i_pol != 2) return fail(nullptr, IRIS_EUNSUP, "i_pol must be 1 or 2");
    if (p->iz_min < 2 || p->iz_min > p->nz) return fail(nullptr, IRIS_EINVAL, "iz_min %d outside [2,nz]", p->iz_min);
  }
  if (p->nof < 1) return fail(nullptr, IRIS_EINVAL, "nof must be >= 1");
  if (p->sample_fourier < 1) return fail(nullptr, IRIS_EINVAL, "sample_fourier must be >= 1");
  if (p->nbox != 1 && p->nbox != 2) return fail(nullptr, IRIS_EINVAL, "nbox must be 1 or 2");
  for (int b = 0; b < p->nbox; ++b) {
    const int32_t* x = p->box[b];
    if (x[0] < 1 || x[0] > x[1] || x[1] > p->nx || x[2] < 1 || x[2] > x[3] || x[3] > p->ny || x[4] < 1 ||
        x[4] > x[5] || x[5] > p->nz)
      return fail(nullptr, IRIS_EINVAL, "Fourier box %d lies outside the interior", b + 1);
  }
  return 0;
}

// Single-step z-chunks of zc planes: lean when every plane z0-1 .. z1+1 lies
// strictly between the UPML slabs, outside the structure and away from the
// TF/SF planes, and the chunk is not the top one; general otherwise.
void single_lists(const iris_params& p, int zc, std::vector<int>& lean, std::vector<int>& gen) {
  auto is_bulk_plane = [&](int z) {
    if (z <= p.pmlzi || z >= p.pmlzf) return false;
    if (z >= p.hi && z <= p.hf) return false;
    if (p.em_source && (z == p.iz_min - 1 || z == p.iz_min)) return false;
    return true;
  };
  lean.clear();
  gen.clear();
  for (int z0 = 1; z0 <= p.nz; z0 += zc) {
    const int z1 = std::min(z0 + zc - 1, p.nz);
    bool ok = z1 < p.nz;
    for (int z = z0 - 1; ok && z <= z1 + 1; ++z) ok = is_bulk_plane(z);
    (ok ? lean : gen).push_back(z0);
  }
}

template <typename T>
void fill_pml(const iris_params& p, const double* fpmn, const double* fmmn, const double* fpmx, const double* fmmx,
              std::vector<PmlPlane<T>>& tab) {
  tab.assign(p.nz + 2, PmlPlane<T>{});
  const double dtq2 = (p.dt / p.q) * (p.dt / p.q);
  for (int iz = 1; iz <= p.nz; ++iz) {
    const double *fp, *fm;
    double ep;
    if (iz <= p.pmlzi) { fp = fpmn + 3 * (iz - 1); fm = fmmn + 3 * (iz - 1); ep = p.eps[0]; }
    else if (iz >= p.pmlzf) { fp = fpmx + 3 * (iz - p.pmlzf); fm = fmmx + 3 * (iz - p.pmlzf); ep = p.eps[2]; }
    else continue;
    PmlPlane<T>& P = tab[iz];
    // component c uses profiles (c+1)%3 for D/B and (c+2)%3 for the E/H decay
    // (reference code/boundary_II.f90:86-99: 1->2,3 ; 2->3,1 ; 3->1,2)
    for (int c = 0; c < 3; ++c) {
      const int pd = (c + 1) % 3, pe = (c + 2) % 3;
      P.fp[c] = (T)fp[c];
      P.fm[c] = (T)fm[c];
      P.cd[c] = (T)(fm[pd] / fp[pd]);
      P.gd[c] = (T)((2.0 * ep) / fp[pd]);
      P.gb[c] = (T)(((2.0 * ep) / fp[pd]) * dtq2);
      P.ce[c] = (T)(fm[pe] / fp[pe]);
      P.ge[c] = (T)(1.0 / (p.eps0 * ep * fp[pe]));
      P.gh[c] = (T)(1.0 / (p.mu0 * fp[pe]));
    }
  }
}

template <typename T>
int upload_tables(iris_handle* h, const double* eps_inv, const double* fpmn, const double* fmmn, const double* fpmx,
                  const double* fmmx) {
  const iris_params& p = h->p;
  // UPML planes
  std::vector<PmlPlane<T>> tab;
  fill_pml<T>(p, fpmn, fmmn, fpmx, fmmx, tab);
  int rc = dalloc(h, &h->pml, sizeof(PmlPlane<T>) * tab.size());
  if (rc) return rc;
  HIP_TRY(h, hipMemcpyAsync(h->pml, tab.data(), sizeof(PmlPlane<T>) * tab.size(), hipMemcpyHostToDevice, h->stream));
  // eps_inv on planes hi..hf (input has planes hi-1..hf+1, x in 0..nx+1)
  const int nzs = p.hf - p.hi + 1;
  const size_t n = (size_t)h->sxy * nzs;
  std::vector<T> re(n, (T)0), im(n, (T)0);
  bool any_im = false;
  const int64_t isx = p.nx + 2, isxy = (int64_t)(p.nx + 2) * (p.ny + 2);
  for (int z = 0; z < nzs; ++z)
    for (int y = 0; y <= p.ny + 1; ++y)
      for (int x = 0; x <= p.nx + 1; ++x) {
        const double* v = eps_inv + 2 * (x + isx * y + isxy * (z + 1));
        const size_t o = (size_t)(x + h->xoff) + (size_t)h->sx * y + (size_t)h->sxy * z;
        re[o] = (T)v[0];
        im[o] = (T)v[1];
        if (v[1] != 0.0) any_im = true;
      }
  rc = dalloc(h, &h->eps_re, sizeof(T) * n);
  if (rc) return rc;
  HIP_TRY(h, hipMemcpyAsync(h->eps_re, re.data(), sizeof(T) * n, hipMemcpyHostToDevice, h->stream));
  if (any_im) {
    rc = dalloc(h, &h->eps_im, sizeof(T) * n);
    if (rc) return rc;
    HIP_TRY(h, hipMemcpyAsync(h->eps_im, im.data(), sizeof(T) * n, hipMemcpyHostToDevice, h->stream));
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

void box_geometry(const iris_params& p, int b, DeviceBox& B) {
  const int32_t* x = p.box[b];
  BoxGeom& g = B.g;
  g.bxi = x[0]; g.bxf = x[1]; g.byi = x[2]; g.byf = x[3]; g.bzi = x[4]; g.bzf = x[5];
  g.nxb = g.bxf + 1 - g.bxi; g.nyb = g.byf + 1 - g.byi; g.nzb = g.bzf + 1 - g.bzi;
  // pairs used by the DFT (reference code/dataprocess.f90:136-207)
  int use[3];
  if (p.nx != 1 && p.ny != 1) { use[0] = use[1] = use[2] = 1; }
  else if (p.ny == 1 && p.nx != 1) { use[0] = 1; use[1] = 0; use[2] = 1; }
  else if (p.nx == 1 && p.ny != 1) { use[0] = 0; use[1] = 1; use[2] = 1; }
  else { use[0] = 0; use[1] = 0; use[2] = 1; }
  const int64_t full[3] = {2LL * 6 * g.nyb * g.nzb, 2LL * 6 * g.nxb * g.nzb, 2LL * 6 * g.nxb * g.nyb};
  B.npts_ref[0] = (p.nx != 1) ? full[0] : 1;
  B.npts_ref[1] = (p.ny != 1) ? full[1] : 1;
  B.npts_ref[2] = full[2];
  int64_t s = 0;
  B.active_mask = 0;
  for (int j = 0; j < 3; ++j) {
    g.npts[j] = use[j] ? full[j] : 0;
    g.start[j] = s;
    s += g.npts[j];
    if (use[j]) B.active_mask |= 1 << j;
  }
  g.ntot = s;
}

// The layout the environment asks for (unset variables: the defaults).
PassConfig pass_config_env(const iris_params& p) {
  PassConfig c;
  auto num = [](const char* name, int* v, int lo) {
    const char* e = std::getenv(name);
    if (e && *e) *v = std::max(lo, std::atoi(e));
  };
  num("IRIS_ZCHUNK", &c.zc, 0);
  if (c.zc > p.nz) c.zc = p.nz;
  num("IRIS_ZCHUNK2", &c.zc2, 0);
  num("IRIS_XCD_ROWS", &c.xcd_rows, 0);
  num("IRIS_XCD_ROWS2", &c.xcd_rows2, 0);
  num("IRIS_TILE_W2", &c.w2, 0);
  num("IRIS_TILE_W2P", &c.w2p, 0);
  num("IRIS_P2_OX", &c.p2ox, 0);
  {
    const char* e = std::getenv("IRIS_NT_STORE2");
    if (e && (e[0] == '0' || e[0] == '1')) c.nt2 = e[0] - '0';
  }
  const char* t = std::getenv("IRIS_TWO_STEP");
  if (t && (t[0] == '0' || t[0] == '1')) c.two = t[0] - '0';
  const char* u = std::getenv("IRIS_TWO_STEP_UPML");
  if (u && (u[0] == '0' || u[0] == '1')) c.upml = u[0] - '0';
  return c;
}

void dfree(iris_handle* h, void** p) {
  if (!*p) return;
  auto it = std::find(h->allocs.begin(), h->allocs.end(), *p);
  if (it != h->allocs.end()) h->allocs.erase(it);
  (void)hipFree(*p);
  *p = nullptr;
}

// A two-step tile of R waves covers footprint cells x0-1 .. x0+W-2 and
// y0-1 .. y0+RY-2 (W = kT2W cells per row, RY = 64 R / W rows); lanes outside
// 1..n stand for the periodic image of one cell inside (a single wrap),
// which holds when the last tile's footprint ends at or before 2n (with
// 64 x 16 tiles: nx >= 32, ny >= 8).
struct T2Geom {
  int W, R, RY, OX, OY;
  int per_cu;  // workgroups of the kernel a CU holds at once
  int tx(int nx) const { return (nx + OX - 1) / OX; }
  int ty(int ny) const { return (ny + OY - 1) / OY; }
};
template <typename T>
T2Geom t2geom(int W, bool pml) {
  T2Geom g;
  g.W = W;
  if (W == 128) {
    // two x-cells per lane (k_yee2p): FP32 fields, plain chunks, one row per wave
    g.R = kP2R;
    g.per_cu = 1;
    g.RY = kP2R;
    g.OX = kP2OX;
    g.OY = kP2OY;
    return g;
  }
  if (W == 32) {
    g.R = pml ? Yee2Shape<T, true, 32>::R : Yee2Shape<T, false, 32>::R;
    g.per_cu = pml ? Yee2Shape<T, true, 32>::WG_PER_CU : Yee2Shape<T, false, 32>::WG_PER_CU;
  } else {
    g.R = pml ? Yee2Shape<T, true, 64>::R : Yee2Shape<T, false, 64>::R;
    g.per_cu = pml ? Yee2Shape<T, true, 64>::WG_PER_CU : Yee2Shape<T, false, 64>::WG_PER_CU;
  }
  g.RY = g.R * kT2X / W;
  g.OX = W - 3;
  g.OY = g.RY - 3;
  return g;
}
// Tiles of the two-cells-a-lane kernel (k_yee2p) over a grid of even nx.
// A tile of ox output cells a row takes rows of sl = ox/2 + 2 lanes, ns rows
// a wave (12 ns footprint rows, 12 ns - 3 output rows).  Main tiles of one
// shape cover x in 1 .. ox tx; narrow tiles (another shape) the r = nx - ox tx
// columns beside them, or all of a grid narrower than a main tile.  The
// main shapes tried fill a wave: ox = 124 (one row a wave), 60 (two), 28
// (four); 60 and 28 output 1260 cells a workgroup against 1116 (fewer halo
// rows: 21 of 24 against 9 of 12).  The layout kept has the fewest
// workgroups a chunk among those whose footprint stays within one period of
// the grid (ties: the wider main tile).
struct P2Layout {
  int ox = 0, sl = 0, ns = 0, tx = 0, ty = 0;             // main tiles
  int oxn = 0, sln = 0, nsn = 0, txn = 0, tyn = 0, x0n = 0;  // narrow tiles
  bool ok = false;
  int64_t tiles() const { return (int64_t)tx * ty + (int64_t)txn * tyn; }
};
// rows of a tile shape that fit the grid's period in y: fewest tile rows
int p2_rows(int ny, int sl, int* ns_out, int* ty_out) {
  for (int ns = 64 / sl; ns >= 1; --ns) {
    const int ry = kP2R * ns, oy = ry - 3;
    const int ty = (ny + oy - 1) / oy;
    if ((int64_t)(ty - 1) * oy + ry - 1 <= 2LL * ny) {
      *ns_out = ns;
      *ty_out = ty;
      return 1;
    }
  }
  return 0;
}
P2Layout p2_layout(int nx, int ny, int only_ox = 0) {
  P2Layout best;
  if (nx < 2 || nx % 2 || ny < 1) return best;
  for (int oxm : {kP2OX, 60, 28}) {
    if (only_ox && oxm != only_ox) continue;
    P2Layout L;
    L.tx = nx / oxm;
    if (L.tx) {
      L.ox = oxm;
      L.sl = oxm / 2 + 2;
      int ns, ty;
      if (!p2_rows(ny, L.sl, &ns, &ty) || ns != 64 / L.sl) continue;  // main tiles fill their waves
      L.ns = ns;
      L.ty = ty;
    }
    const int r = nx - oxm * L.tx;
    if (r) {
      int64_t bn = -1;
      for (int m = 1; m <= 8; ++m) {
        const int ox = ((r + m - 1) / m + 1) / 2 * 2;
        const int sl = ox / 2 + 2;
        if (sl > 64) continue;
        const int txn = (r + ox - 1) / ox;
        if ((int64_t)oxm * L.tx + (int64_t)txn * ox + 2 > 2LL * nx) continue;
        int ns, tyn;
        if (!p2_rows(ny, sl, &ns, &tyn)) continue;
        if (bn < 0 || (int64_t)txn * tyn < bn) {
          bn = (int64_t)txn * tyn;
          L.oxn = ox;
          L.sln = sl;
          L.nsn = ns;
          L.txn = txn;
          L.tyn = tyn;
        }
      }
      if (bn < 0) continue;
      L.x0n = oxm * L.tx;
    }
    L.ok = true;
    if (!best.ok || L.tiles() < best.tiles()) best = L;
  }
  return best;
}

bool two_step_fits(const iris_params& p, const T2Geom& g) {
  return (int64_t)(g.tx(p.nx) - 1) * g.OX + g.W - 1 <= 2LL * p.nx &&
         (int64_t)(g.ty(p.ny) - 1) * g.OY + g.RY - 1 <= 2LL * p.ny;
}

// (Re)builds the chunk lists of the step passes for layout c.  The field
// state is not touched; the handle's stream is synchronised.
template <typename T>
int build_passes(iris_handle* h, const PassConfig& c) {
  const iris_params& p = h->p;
  int rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  graph_release(h);  // captured with the old lists
  ++h->pass_gen;     // a batch made from the old lists refuses to step
  h->pass = c;
  h->xcd_rows = c.xcd_rows;
  h->xcd_rows2 = c.xcd_rows2 >= 0 ? c.xcd_rows2 : 4;  // resolved below once the tile width is known
  // Streaming stores of the two-step outputs: measured 6-10% faster at
  // 128 x 128 x 768 and, with FP32 fields one cell a lane, 3-5% slower at
  // 512 x 512 x 1024 (same box); on for two cells a lane below
  h->nt2 = c.nt2 >= 0 ? c.nt2 : (p.nx <= 160 && p.ny <= 160 ? 1 : 0);
  const size_t fbytes = h->csize * 3 * (size_t)h->cs;
  const size_t abytes = h->csize * 3 * (size_t)(h->ps0 + h->ps1);
  {
    const int fgx = (p.nx + kFTX - 1) / kFTX, fgy = (p.ny + kFTY - 1) / kFTY;
    int fzc = (int)std::min<int64_t>(32, std::max<int64_t>(1, (int64_t)fgx * fgy * p.nz / 4096));
    if (c.zc > 0) fzc = std::min(c.zc, p.nz);
    h->grd_fused = dim3(fgx, fgy, (p.nz + fzc - 1) / fzc);
    if (h->fused) h->zchunk = fzc;
  }
  // no two-step pass until the lists below say otherwise
  h->two_step = false;
  h->n2 = h->n2p = 0;
  h->planes2 = h->planes2p = h->upml2p = 0;
  h->rest_groups.clear();
  h->nrest_lean = h->nrest_gen = 0;
  for (int k = 0; k < 2; ++k) {
    dfree(h, (void**)&h->rest_wg[k]);
    h->rest_nwg[k] = 0;
  }
  for (auto& t : h->rest_tabs) dfree(h, &t.dev);
  h->rest_tabs.clear();
  // UPML-variant chunks: up to 2 zc planes; plain two-step chunks: below
  const int zc2p = std::min(2 * h->zchunk, kChunkLenMask + 1);
  int zc2 = zc2p;
  {
    // fused z-chunks: a chunk is bulk when every plane it reads or writes,
    // z0-1 .. z1+1, lies strictly between the UPML slabs, outside the
    // structure and away from the TF/SF planes, and it is not the top chunk
    std::vector<int> bulk, gen;
    const int zc = h->zchunk;
    single_lists(p, zc, bulk, gen);
    dfree(h, (void**)&h->chunks);
    auto is_bulk_plane = [&](int z) {
      if (z <= p.pmlzi || z >= p.pmlzf) return false;
      if (z >= p.hi && z <= p.hf) return false;
      if (p.em_source && (z == p.iz_min - 1 || z == p.iz_min)) return false;
      return true;
    };
    h->nbulk = (int)bulk.size();
    h->ngen = (int)gen.size();
    std::vector<int> all(bulk);
    all.insert(all.end(), gen.begin(), gen.end());
    if ((rc = dalloc(h, (void**)&h->chunks, sizeof(int) * all.size()))) return rc;
    HIP_TRY(h, hipMemcpyAsync(h->chunks, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));

    // Two steps per pass.  An output plane z of a two-step chunk needs
    // 2 <= z <= nz-2 (the wrap in z stays with single steps) and no TF/SF plane
    // in z-1 .. z+2 (the corrections belong to single steps); the plain kernel
    // also needs z-1 .. z+2 outside the UPML slabs, the others take its UPML
    // variant.  Runs of planes of one kind are cut into near-equal chunks of
    // up to 2*zc planes; the remaining planes take single
    // steps in chunks of up to zc planes with their own lengths.
    // c.two = 0: off, 1: on wherever it applies, -1: on when its tiles
    // waste fewer lanes than the single-step pass's, as estimated by
    // the lanes per cell-step of both (64 x R footprint lanes per 61 x (R-3)
    // cells and two steps, against 64 x 9 lanes per 64 x 8 cells and one
    // step).  Measured on MI355X with 768-plane grids, the two-step pass is
    // slower only where this ratio is above 1 (one 64-wide row of tiles
    // against two).
    // tile width: 64 cells a row by default, or 32 (two rows per wave;
    // iris_autotune tries it); the other where only that one fits.  With
    // FP32 fields the 32-wide tiles (8 waves, two workgroups a CU) measured
    // within noise of the 64-wide ones at 512 x 512 x 1024 and 12% faster at
    // 128 x 128 x 768; on the 64-grating sweep, batched, 3% slower.
    // 128: two x-cells per lane (k_yee2p), with FP32 fields on grids of even
    // nx only (a lane's pair of cells must not straddle the x period), and
    // the default there: main tiles of 124, 60 or 28 cells a row, the columns
    // beside (and grids narrower than those) in narrow tiles of the same
    // launch (p2_layout).  With 124-cell tiles alone, measured at ny = 512,
    // nz = 1024 against 64-wide tiles: 21% faster at nx = 64, 13% at 100, 5%
    // at 128, 13% at 250 and 256, 8% at 376, 22-28% at 496, 17% at 512, 15%
    // at 1024.  At 512 x 512 x 1024 the 60-cell main tiles (216 workgroups a
    // chunk against 236) measured within noise of the 124-cell ones, 28-cell
    // ones 15% slower.
    P2Layout pl = p2_layout(p.nx, p.ny, c.p2ox);
    if (!pl.ok && c.p2ox) pl = p2_layout(p.nx, p.ny);
    const bool pair_ok = sizeof(T) == 4 && pl.ok;
    const bool pair_default = pair_ok;
    int W = (c.w2 == 32 || c.w2 == 64 || (c.w2 == 128 && pair_ok)) ? c.w2 : (c.w2 == 0 && pair_default ? 128 : 64);
    if (W != 128 && !two_step_fits(p, t2geom<T>(W, false)) && two_step_fits(p, t2geom<T>(96 - W, false)))
      W = 96 - W;
    h->t2w = W;
    // streaming stores at 512 x 512 x 1024 with two cells a lane: within
    // noise, ~2% faster per launch (four runs of each in ABBA order; with
    // FP64 fields ~2% slower, so those keep the rule above)
    if (c.nt2 < 0 && W == 128) h->nt2 = 1;
    // tile rows per XCD run: 2 with two cells a lane (measured 3% faster
    // than 4 at 512 x 512 x 1024 over four interleaved runs each)
    if (c.xcd_rows2 < 0) h->xcd_rows2 = W == 128 ? 2 : 4;
    // the UPML variant's own width, 32 by default: its slabs' launch is a
    // few rounds of workgroups, and with FP32 fields 32-wide tiles of 12
    // waves cover a 512 x 512 plane in 450 workgroups against 513 (two slabs:
    // 4 rounds on 256 CUs instead of 5).  Measured at 512 x 512 x 1024: 13-16%
    // faster with FP32 fields, 14% with FP64.
    int Wp = (c.w2p == 32 || c.w2p == 64) ? c.w2p : 32;
    if (!two_step_fits(p, t2geom<T>(Wp, true)) && two_step_fits(p, t2geom<T>(96 - Wp, true))) Wp = 96 - Wp;
    h->t2wp = Wp;
    const T2Geom g2 = t2geom<T>(W, false), g2p = t2geom<T>(Wp, true);
    // two-step tiles a chunk of planes
    const int64_t t2 = W == 128 ? pl.tiles() : (int64_t)g2.tx(p.nx) * g2.ty(p.ny);
    bool want = c.two != 0;
    if (c.two < 0) {
      const double two = (double)t2 * g2.R / 2.0;
      const double one = (double)((p.nx + kFTX - 1) / kFTX) * ((p.ny + kFTY - 1) / kFTY) * (kFTY + 1);
      if (two > 0.95 * one) want = false;
    }
    h->bloch = !(p.ak[0][0] == 0.0 && p.ak[0][1] == 0.0 && p.ak[1][0] == 0.0 && p.ak[1][1] == 0.0);
    const bool shape = h->fused && want && (W == 128 ? pl.ok : two_step_fits(p, g2)) && p.nz >= 6;
    const bool shape_pml = two_step_fits(p, g2p);
    auto tfsf_plane = [&](int z) { return p.em_source && (z == p.iz_min - 1 || z == p.iz_min); };
    auto upml_plane = [&](int z) { return z <= p.pmlzi || z >= p.pmlzf; };
    // 0: single step; 1: plain two-step output plane; 2: UPML-variant output plane
    std::vector<char> kind((size_t)p.nz + 2, 0);
    int nout = 0;
    for (int z = 2; shape && z <= p.nz - 2; ++z) {
      bool ok = true;
      for (int q = z - 1; ok && q <= z + 2; ++q) ok = !tfsf_plane(q);
      if (!ok) continue;
      bool plain = true;
      for (int q = z - 1; plain && q <= z + 2; ++q) plain = !upml_plane(q);
      if (!plain && !shape_pml) continue;
      kind[z] = plain ? 1 : 2;
      ++nout;
    }
    // The UPML variant by default only where its launch gives at least every
    // other CU a workgroup: its slabs are a few dozen planes, and on small
    // grids the few workgroups of its launch took longer than the single
    // steps they replace (measured on 64 gratings of 32..128 x 32..128 x 768
    // cells).  With FP32 fields and a Bloch phase on n x n x 768 grids it
    // measured 7% faster a step at n = 192 (140 workgroups) and 224 (176),
    // 6% on 256 x 256 x 1024 (234), within 1% at n = 128 and 160 (70, 96).
    // c.upml = 0: never, 1: wherever it applies, -1: by that rule.
    bool upml_two = c.upml != 0;
    if (c.upml < 0) {
      int64_t nch = 0;
      for (int z = 1; z <= p.nz; ++z) {
        if (kind[z] != 2) continue;
        int e = z;
        while (e + 1 <= p.nz && kind[e + 1] == 2) ++e;
        nch += (e - z + 1 + zc2p - 1) / zc2p;
        z = e;
      }
      const int64_t tiles = (int64_t)g2p.tx(p.nx) * g2p.ty(p.ny);
      if (2 * tiles * nch < h->ncu) upml_two = false;
    }
    if (!upml_two)
      for (int z = 1; z <= p.nz; ++z)
        if (kind[z] == 2) { kind[z] = 0; --nout; }
    // a run of two-step planes needs at least two of them of each kind
    for (int z = 1; z <= p.nz; ++z) {
      if (!kind[z]) continue;
      int e = z;
      while (e + 1 <= p.nz && kind[e + 1] == kind[z]) ++e;
      if (e == z) { kind[z] = 0; --nout; }
      z = e;
    }
    // Plain chunk length.  One workgroup of the plain kernel fills a CU (16
    // waves with FP32 fields, 12 with FP64, and most of the LDS), so a launch
    // takes ceil(workgroups / CUs) rounds of (planes + 3) pipeline steps.
    // When chunks of 2 zc planes take few rounds, a nearly empty last round
    // costs as much as a full one (on a 128 x 128 x 768 grid 1830
    // workgroups, 8 rounds, ran 7% slower than 1770, 7 rounds); then the
    // length is the one of zc .. 4 zc with the fewest such steps (the
    // shortest among equals): 22 planes there, 5-7% faster.  With many
    // rounds the last one matters little and longer chunks measured slower
    // (82 against 61 planes on 512 x 512 x 1024: 3-6%), so 2 zc stays;
    // there 24, 32, 39 and 48 planes with two cells a lane measured within
    // noise of 64 (four interleaved runs each).
    auto plain_runs = [&]() {
      std::vector<int> runs;
      for (int z = 1; z <= p.nz; ++z) {
        if (kind[z] != 1) continue;
        int e = z;
        while (e + 1 <= p.nz && kind[e + 1] == 1) ++e;
        runs.push_back(e - z + 1);
        z = e;
      }
      return runs;
    };
    auto rounds = [&](const std::vector<int>& runs, int L) {
      int64_t nch = 0;
      for (int r : runs) nch += (r + L - 1) / L;
      const int64_t slots = (int64_t)h->ncu * g2.per_cu;
      return (t2 * nch + slots - 1) / slots;
    };
    if (c.zc2 > 0) {
      zc2 = std::min(c.zc2, kChunkLenMask + 1);
    } else if (nout > 0 && rounds(plain_runs(), zc2p) < 16) {
      const std::vector<int> runs = plain_runs();
      int64_t best = -1;
      for (int L = h->zchunk; L <= std::min(4 * h->zchunk, kChunkLenMask + 1); ++L) {
        const int64_t cost = rounds(runs, L) * (L + 3);
        if (best < 0 || cost < best) {
          best = cost;
          zc2 = L;
        }
      }
    }
    h->zc2 = zc2;
    if (nout > 0) {
      std::vector<int> c2, c2p;
      std::map<int, std::pair<std::vector<int>, std::vector<int>>> byl;  // length -> (lean, general)
      std::vector<unsigned char> zm((size_t)p.nz + 2, 0);
      for (int z = 1; z <= p.nz;) {
        if (!kind[z]) {
          // single-step planes z .. e: chunks of zc planes and one shorter
          int e = z;
          while (e + 1 <= p.nz && !kind[e + 1]) ++e;
          for (int c0 = z; c0 <= e; c0 += zc) {
            const int c1 = std::min(c0 + zc - 1, e);
            bool lean = c1 < p.nz;
            for (int q = c0 - 1; lean && q <= c1 + 1; ++q) lean = is_bulk_plane(q);
            auto& g = byl[c1 - c0 + 1];
            (lean ? g.first : g.second).push_back(c0);
          }
          z = e + 1;
          continue;
        }
        // a run of two-step planes z .. e (both kinds); exports at its ends
        int e = z;
        while (e + 1 <= p.nz && kind[e + 1]) ++e;
        for (int s0 = z; s0 <= e;) {
          int s1 = s0;
          while (s1 + 1 <= e && kind[s1 + 1] == kind[s0]) ++s1;
          // near-equal chunks of at most zc2 planes (a short last chunk
          // would repeat the pipeline's 3 planes of fill for few planes)
          const int step = kind[s0] == 1 ? zc2 : zc2p;
          const int nseg = s1 - s0 + 1, nch = (nseg + step - 1) / step;
          for (int k = 0, c0 = s0; k < nch; ++k) {
            const int len = nseg / nch + (k < nseg % nch ? 1 : 0);
            int v = c0 | ((len - 1) << kChunkLenShift);
            if (c0 == z) v |= kChunkExportBelow;
            if (c0 + len - 1 == e) v |= kChunkExportAbove;
            (kind[s0] == 1 ? c2 : c2p).push_back(v);
            c0 += len;
          }
          for (int q = s0; q <= s1; ++q) {
            zm[q] = 1;
            if (kind[s0] == 1) h->planes2 += 1;
            else { h->planes2p += 1; if (upml_plane(q)) h->upml2p += 1; }
          }
          s0 = s1 + 1;
        }
        z = e + 1;
      }
      std::vector<int> rest;
      for (auto& kv : byl) {
        RestGroup G{kv.first, (int)rest.size(), (int)kv.second.first.size(), (int)kv.second.second.size()};
        rest.insert(rest.end(), kv.second.first.begin(), kv.second.first.end());
        rest.insert(rest.end(), kv.second.second.begin(), kv.second.second.end());
        h->rest_groups.push_back(G);
        h->nrest_lean += G.nlean;
        h->nrest_gen += G.ngen;
      }
      if (h->rest_groups.size() > 1) {
        const int t = (int)(h->grd_fused.x * h->grd_fused.y);
        std::vector<int> w[2] = {{0}, {0}};
        for (const RestGroup& G : h->rest_groups) {
          w[0].push_back(w[0].back() + t * G.nlean);
          w[1].push_back(w[1].back() + t * G.ngen);
        }
        for (int k = 0; k < 2; ++k) {
          if ((rc = dalloc(h, (void**)&h->rest_wg[k], sizeof(int) * w[k].size()))) return rc;
          HIP_TRY(h, hipMemcpyAsync(h->rest_wg[k], w[k].data(), sizeof(int) * w[k].size(), hipMemcpyHostToDevice,
                                    h->stream));
          h->rest_nwg[k] = w[k].back();
        }
        HIP_TRY(h, hipStreamSynchronize(h->stream));
      }
      h->two_step = true;
      h->n2 = (int)c2.size();
      h->n2p = (int)c2p.size();
      h->tx2 = g2.tx(p.nx);
      h->ty2 = g2.ty(p.ny);
      h->tx2p = g2p.tx(p.nx);
      h->ty2p = g2p.ty(p.ny);
      h->p2ox = h->p2sl = h->p2ns = h->p2oxn = h->p2sln = h->p2nsn = h->tx2n = h->ty2n = h->x0n = 0;
      if (W == 128) {
        h->tx2 = pl.tx;
        h->ty2 = pl.ty;
        h->p2ox = pl.ox;
        h->p2sl = pl.sl;
        h->p2ns = pl.ns;
        h->p2oxn = pl.oxn;
        h->p2sln = pl.sln;
        h->p2nsn = pl.nsn;
        h->tx2n = pl.txn;
        h->ty2n = pl.tyn;
        h->x0n = pl.x0n;
      }
      dfree(h, (void**)&h->chunks2);
      dfree(h, (void**)&h->chunks2p);
      dfree(h, (void**)&h->chunks_rest);
      dfree(h, (void**)&h->zmask);
      if ((rc = dalloc(h, (void**)&h->chunks2, sizeof(int) * std::max<size_t>(c2.size(), 1))) ||
          (rc = dalloc(h, (void**)&h->chunks2p, sizeof(int) * std::max<size_t>(c2p.size(), 1))) ||
          (rc = dalloc(h, (void**)&h->chunks_rest, sizeof(int) * std::max<size_t>(rest.size(), 1))) ||
          (rc = dalloc(h, (void**)&h->zmask, zm.size())))
        return rc;
      if (!h->ke && ((rc = dalloc(h, &h->ke, fbytes)) || (rc = dalloc(h, &h->kh, fbytes)))) return rc;
      if (!h->kd && ((rc = dalloc(h, &h->kd, abytes)) || (rc = dalloc(h, &h->kb, abytes)))) return rc;
      if (!c2.empty())
        HIP_TRY(h, hipMemcpyAsync(h->chunks2, c2.data(), sizeof(int) * c2.size(), hipMemcpyHostToDevice, h->stream));
      if (!c2p.empty())
        HIP_TRY(h, hipMemcpyAsync(h->chunks2p, c2p.data(), sizeof(int) * c2p.size(), hipMemcpyHostToDevice,
                                  h->stream));
      if (!rest.empty())
        HIP_TRY(h, hipMemcpyAsync(h->chunks_rest, rest.data(), sizeof(int) * rest.size(), hipMemcpyHostToDevice,
                                  h->stream));
      HIP_TRY(h, hipMemcpyAsync(h->zmask, zm.data(), zm.size(), hipMemcpyHostToDevice, h->stream));
      HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

template <typename T>
int create_impl(iris_handle* h, const double* eps_inv, const double* fpmn, const double* fmmn, const double* fpmx,
                const double* fmmx, const double* fw) {
  const iris_params& p = h->p;
  h->csize = sizeof(cplx<T>);
  // x pitch: ix=1 on a 128-byte boundary
  const int align = (int)(128 / sizeof(cplx<T>));
  h->xoff = align - 1;
  h->sx = ((h->xoff + p.nx + 2 + align - 1) / align) * align;
  h->sy = p.ny + 2;
  h->sxy = h->sx * h->sy;
  h->ncell = h->sxy * (p.nz + 2);
  int rc;
  // component stride rounded to 256 elements keeps every component aligned
  h->cs = ((h->ncell + 255) / 256) * 256;
  const size_t fbytes = h->csize * 3 * (size_t)h->cs;
  if ((rc = dalloc(h, &h->fe, fbytes)) || (rc = dalloc(h, &h->fh, fbytes))) return rc;
  if (h->fused && ((rc = dalloc(h, &h->ge, fbytes)) || (rc = dalloc(h, &h->gh, fbytes)))) return rc;
  h->ps0 = h->sxy * p.pmlzi;
  h->ps1 = h->sxy * (p.nz - p.pmlzf + 1);
  h->pb1 = 3 * h->ps0;
  const size_t abytes = h->csize * 3 * (size_t)(h->ps0 + h->ps1);
  if ((rc = dalloc(h, &h->d, abytes)) || (rc = dalloc(h, &h->b, abytes))) return rc;
  if (h->fused && ((rc = dalloc(h, &h->dn, abytes)) || (rc = dalloc(h, &h->bn, abytes)))) return rc;
  if ((rc = upload_tables<T>(h, eps_inv, fpmn, fmmn, fpmx, fmmx))) return rc;
  // DFT
  if ((rc = dalloc(h, (void**)&h->fw, sizeof(double) * p.nof))) return rc;
  HIP_TRY(h, hipMemcpyAsync(h->fw, fw, sizeof(double) * p.nof, hipMemcpyHostToDevice, h->stream));
  if ((rc = dalloc(h, (void**)&h->ph, sizeof(double2) * (size_t)p.nof * dft_slots<T>()))) return rc;
  if ((rc = dalloc(h, (void**)&h->flux, sizeof(double) * 6 * (size_t)p.nof))) return rc;
  for (int b = 0; b < p.nbox; ++b) {
    DeviceBox& B = h->box[b];
    box_geometry(p, b, B);
    if ((rc = dalloc(h, (void**)&B.emg, sizeof(double2) * (size_t)std::max<int64_t>(B.g.ntot, 1) * p.nof))) return rc;
    if ((rc = dalloc(h, &B.samp, h->csize * (size_t)std::max<int64_t>(B.g.ntot, 1) * dft_slots<T>()))) return rc;
  }
  // launch shapes: x-rows of up to 64 (one wavefront), 256 threads per block
  const int bx = std::min(64, pow2ceil(p.nx));
  const int by = std::max(1, std::min(256 / bx, pow2ceil(p.ny)));
  h->blk_yee = dim3(bx, by, 1);
  const int gx = (p.nx + bx - 1) / bx, gy = (p.ny + by - 1) / by;
  const int64_t nbxy = (int64_t)gx * gy;
  int zc = (int)std::min<int64_t>(64, std::max<int64_t>(1, nbxy * p.nz / 4096));
  h->zchunk = zc;
  h->grd_yee = dim3(gx, gy, (p.nz + zc - 1) / zc);
  h->blk_fused = dim3(kFTX, kFTY + 1, 1);
  const int tbx = std::min(64, pow2ceil(p.nx + 1));
  const int tby = std::max(1, std::min(256 / tbx, pow2ceil(p.ny + 1)));
  h->blk_tf = dim3(tbx, tby, 1);
  h->grd_tf = dim3((p.nx + 1 + tbx - 1) / tbx, (p.ny + 1 + tby - 1) / tby, 1);
  h->dtcq2 = ((p.dt * p.c0) / p.q) * ((p.dt * p.c0) / p.q);
  {
    TfsfConst t = tfsf_const(h);
    if (incident_matched(p)) {
      // off normal incidence: the incident field matched to the Bloch condition
      if (const char* why = incident_unsupported(p)) return fail(h, IRIS_EUNSUP, "%s", why);
      const std::vector<zcplx> tab = incident_table(p);
      void* dtab = nullptr;
      if ((rc = dalloc(h, &dtab, sizeof(zcplx) * tab.size()))) return rc;
      // on the handle's stream, after dalloc's zero fill there (a copy on the
      // null stream is not ordered with this non-blocking stream)
      HIP_TRY(h, hipMemcpyAsync(dtab, tab.data(), sizeof(zcplx) * tab.size(), hipMemcpyHostToDevice, h->stream));
      HIP_TRY(h, hipStreamSynchronize(h->stream));  // `tab` is freed at the end of this block
      t.tab = (const double2*)dtab;
      t.phix = p.ak[0][0] / p.nx;
      t.mode2d = 0;  // the corrections of 3D grids (exact for a discrete wave, also on 1D/2D grids)
    }
    if ((rc = dalloc(h, (void**)&h->tfc, sizeof(TfsfConst)))) return rc;
    HIP_TRY(h, hipMemcpyAsync(h->tfc, &t, sizeof t, hipMemcpyHostToDevice, h->stream));
  }
  if ((rc = build_passes<T>(h, pass_config_env(p)))) return rc;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

// ---- gradient: snapshots, recomputation, adjoint steps ---------------------

size_t field_bytes(const iris_handle* h) { return h->csize * 3 * (size_t)h->cs; }
size_t aux_bytes(const iris_handle* h) { return h->csize * 3 * (size_t)(h->ps0 + h->ps1); }
size_t state_bytes(const iris_handle* h) { return 2 * field_bytes(h) + 2 * aux_bytes(h); }

// copy the current state (E, H, D, B) into a snapshot buffer, or back
int state_copy(iris_handle* h, void* snap, bool save) {
  ProfScope ps(h, IRIS_K_CHECKPOINT);
  const size_t fb = field_bytes(h), ab = aux_bytes(h);
  char* s = (char*)snap;
  void* parts[4] = {h->fe, h->fh, h->d, h->b};
  const size_t off[4] = {0, fb, 2 * fb, 2 * fb + ab};
  const size_t len[4] = {fb, fb, ab, ab};
  for (int k = 0; k < 4; ++k) {
    if (save) HIP_TRY(h, hipMemcpyAsync(s + off[k], parts[k], len[k], hipMemcpyDeviceToDevice, h->stream));
    else HIP_TRY(h, hipMemcpyAsync(parts[k], s + off[k], len[k], hipMemcpyDeviceToDevice, h->stream));
  }
  return 0;
}

int adj_alloc(iris_handle* h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(h, IRIS_ENOMEM, "hipMalloc(%zu bytes) for the adjoint failed: %s", bytes, hipGetErrorString(e));
  }
  h->adj_allocs.push_back(*p);
  return 0;
}

template <typename T>
int run_tail(iris_handle* h, int it_first, int it_last);

// iris_step with snapshots: before tail_t0 the steps are cut at the snapshot
// boundaries so that the state after step k*ck_seg is whole when it is
// copied; from tail_t0 on H of every step is recorded
int step_with_checkpoints(iris_handle* h, int it_first, int it_last) {
  int it = it_first;
  while (it <= it_last) {
    const int cur = it - 1;
    int rc;
    if (cur < h->tail_t0) {
      const int k = cur / h->ck_seg;
      if (cur % h->ck_seg == 0 && k < (int)h->ck.size() && !h->ck_ok[k]) {
        if ((rc = state_copy(h, h->ck[k], true))) return rc;
        h->ck_ok[k] = 1;
      }
      const int end = std::min(it_last, std::min((k + 1) * h->ck_seg, h->tail_t0));
      rc = h->prec == 8 ? run_steps<double>(h, it, end) : run_steps<float>(h, it, end);
      if (rc) return rc;
      it = end + 1;
    } else {
      rc = h->prec == 8 ? run_tail<double>(h, it, it_last) : run_tail<float>(h, it, it_last);
      if (rc) return rc;
      it = it_last + 1;
    }
  }
  return 0;
}

size_t rec_step_bytes(const iris_handle* h) { return h->csize * 3 * (size_t)h->rec_plane; }
const void* rec_at(const iris_handle* h, const void* base, int r) {
  return (const char*)base + rec_step_bytes(h) * (size_t)r;
}

// H on planes hi-1..hf of field buffer f into slot r of record buffer base
int record_h(iris_handle* h, const void* f, void* base, int r) {
  ProfScope ps(h, IRIS_K_CHECKPOINT);
  // component strides and the plane range are multiples of 16 bytes (x pitch
  // a multiple of 128 bytes, component stride of 256 elements)
  const int64_t n16 = (int64_t)(h->csize * (size_t)h->rec_plane / 16);
  const int64_t cs16 = (int64_t)(h->csize * (size_t)h->cs / 16);
  const char* src = (const char*)f + h->csize * (size_t)h->sxy * (h->p.hi - 1);
  char* dst = (char*)base + rec_step_bytes(h) * (size_t)r;
  hipLaunchKernelGGL(k_record, dim3((unsigned)((3 * n16 + 255) / 256)), dim3(256), 0, h->stream, (const int4*)src,
                     cs16, n16, (int4*)dst);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

// One step of the adjoint, the transpose of step n (the operations of the
// forward step in reverse order; the TF/SF terms do not depend on the state).
// face_on bit 2j+s: face half s of pair j carries objective weight.
template <typename T>
int adjoint_step(iris_handle* h, int n, int nobox, int face_on, const void* rec) {
  using C = cplx<T>;
  const iris_params& p = h->p;
  ProfScope ps(h, IRIS_K_ADJOINT);
  StepArgs<T> a = step_args<T>(h);
  C* ae = (C*)h->ae;
  C* ah = (C*)h->ah;
  C* tmp = (C*)h->atmp;
  const int64_t ncells = (int64_t)p.nx * p.ny * p.nz;
  const int64_t ncells1 = (int64_t)(p.nx + 1) * (p.ny + 1) * (p.nz + 1);
  // 1. sources: the DFT sample taken after step n
  if (n >= p.to_fourier && (n % p.sample_fourier) == 0 && face_on) {
    hipLaunchKernelGGL(k_dft_phase, dim3((p.nof + 255) / 256), dim3(256), 0, h->stream, h->fw, p.nof, 1, 1,
                       n + 1 - p.to_fourier, p.sample_fourier, 1.0, h->ph, nullptr);
    const DeviceBox& B = h->box[nobox - 1];
    for (int j = 0; j < 3; ++j)
      for (int s = 0; s < 2; ++s) {
        if (!((face_on >> (2 * j + s)) & 1)) continue;
        hipLaunchKernelGGL(k_adj_source<T>, dim3(blocks_for(B.g.npts[j] / 2)), dim3(256), 0, h->stream, B.g, j, s,
                           (const double2*)h->adj_g, p.nof, (const double2*)h->ph, p.dt, ae, ah, h->cs, h->sx,
                           h->sxy, h->xoff);
      }
  }
  // 2. min-side wraps of H (ghost 0 <- exp(-i ak) H(n))
  const int nn[3] = {p.nx, p.ny, p.nz};
  const C fm[3] = {a.fxm, a.fym, a.fzm}, fp[3] = {a.fxp, a.fyp, a.fzp};
  for (int ax = 0; ax < 3; ++ax) {
    const int na = (ax == 0) ? p.ny + 2 : p.nx + 2, nb = (ax == 2) ? p.ny + 2 : p.nz + 2;
    hipLaunchKernelGGL(k_adj_wrap<T>, dim3(blocks_for(3LL * na * nb)), dim3(256), 0, h->stream, ah, h->cs, ax, 0,
                       nn[ax], fm[ax], p.nx, p.ny, p.nz, h->sx, h->sxy, h->xoff);
  }
  // 3. H update: local part, then E^ += curl_b(.)
  a.d = (C*)h->ad;
  a.b = (C*)h->ab;
  hipLaunchKernelGGL(k_adj_h_local<T>, dim3(blocks_for(ncells)), dim3(256), 0, h->stream, a, ah, (C*)h->ab, tmp);
  hipLaunchKernelGGL(k_adj_curl_b<T>, dim3(blocks_for(ncells1)), dim3(256), 0, h->stream, a, (const C*)tmp, ae);
  // 4. max-side wraps of E (ghost n+1 <- exp(+i ak) E(1))
  for (int ax = 0; ax < 3; ++ax) {
    const int na = (ax == 0) ? p.ny + 2 : p.nx + 2, nb = (ax == 2) ? p.ny + 2 : p.nz + 2;
    hipLaunchKernelGGL(k_adj_wrap<T>, dim3(blocks_for(3LL * na * nb)), dim3(256), 0, h->stream, ae, h->cs, ax,
                       nn[ax] + 1, 1, fp[ax], p.nx, p.ny, p.nz, h->sx, h->sxy, h->xoff);
  }
  // 5. gradient on the structure cells with H of step n-1
  const int64_t nstruct = (int64_t)p.nx * p.ny * (p.hf - p.hi + 1);
  hipLaunchKernelGGL(k_adj_grad<T>, dim3(blocks_for(nstruct)), dim3(256), 0, h->stream, a, (const C*)ae,
                     (const C*)rec, h->rec_plane, h->grad);
  // 6. E update: local part, then H^ += curl_f(.)
  hipLaunchKernelGGL(k_adj_e_local<T>, dim3(blocks_for(ncells)), dim3(256), 0, h->stream, a, ae, (C*)h->ad, tmp);
  hipLaunchKernelGGL(k_adj_curl_f<T>, dim3(blocks_for(ncells1)), dim3(256), 0, h->stream, a, (const C*)tmp, ah);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

// StepArgs of the fused adjoint passes: the forward geometry and
// coefficients with the adjoint's Bloch factors, conj(exp(-+i ak)) in place
// of exp(+-i ak) (fxm forms the min-side ghosts of Y, fxp the max-side
// images of X)
template <typename T>
StepArgs<T> adj_args(const iris_handle* h, void* e, void* hf, void* en, void* hn, void* d, void* b, void* dn,
                     void* bn) {
  StepArgs<T> a = step_args_on<T>(h, e, hf, en, hn, d, b, dn, bn);
  const cplx<T> xp = a.fxp, yp = a.fyp, zp = a.fzp;
  a.fxp = cconj(a.fxm); a.fyp = cconj(a.fym); a.fzp = cconj(a.fzm);
  a.fxm = cconj(xp); a.fym = cconj(yp); a.fzm = cconj(zp);
  return a;
}

// objective sources of step n into the adjoint state held by a (e, h)
template <typename T>
void adj_sources(iris_handle* h, const StepArgs<T>& a, int n, int nobox, int face_on) {
  const iris_params& p = h->p;
  if (!(n >= p.to_fourier && (n % p.sample_fourier) == 0 && face_on)) return;
  hipLaunchKernelGGL(k_dft_phase, dim3((p.nof + 255) / 256), dim3(256), 0, h->stream, h->fw, p.nof, 1, 1,
                     n + 1 - p.to_fourier, p.sample_fourier, 1.0, h->ph, nullptr);
  const DeviceBox& B = h->box[nobox - 1];
  const int lo[3] = {B.g.bxi, B.g.byi, B.g.bzi}, hi[3] = {B.g.bxf, B.g.byf, B.g.bzf};
  for (int j = 0; j < 3; ++j) {
    const int m = (face_on >> (2 * j)) & 3;
    if (!m) continue;
    if (m == 3 && lo[j] != hi[j]) {
      // the two faces of the pair share no cell: one launch
      hipLaunchKernelGGL(k_adj_source_mixed<T>, dim3(blocks_for(B.g.npts[j])), dim3(256), 0, h->stream, a, B.g, j, 2,
                         (const double2*)h->adj_g, p.nof, (const double2*)h->ph, p.dt);
      continue;
    }
    for (int s = 0; s < 2; ++s)
      if ((m >> s) & 1)
        hipLaunchKernelGGL(k_adj_source_mixed<T>, dim3(blocks_for(B.g.npts[j] / 2)), dim3(256), 0, h->stream, a,
                           B.g, j, s, (const double2*)h->adj_g, p.nof, (const double2*)h->ph, p.dt);
  }
}

// one fused adjoint step over lists of lean chunks (the forward's lean pass
// on the plain planes) and general chunks (k_adj_gen)
template <typename T>
void adj_lists(iris_handle* h, const StepArgs<T>& a, const int* lean, int nlean, const int* gen, int ngen,
               const void* recp) {
  using C = cplx<T>;
  const C* rec = (const C*)recp;
  TileMap tm{(int)h->grd_fused.x, (int)h->grd_fused.y, 0, (int)h->grd_fused.x * h->xcd_rows};
  if (nlean) {
    tm.nblk = tm.tx * tm.ty * nlean;
    hipLaunchKernelGGL((k_yee_fused<T, false>), dim3(tm.nblk), h->blk_fused, 0, h->stream, a,
                       (const TfsfConst*)h->tfc, 0, lean, tm);
  }
  if (ngen) {
    // tile rows of the general adjoint pass: 7 rows + the halo row make 8
    // waves, two workgroups per CU at the FP32 kernel's 122 VGPRs (tuning:
    // IRIS_ADJ_TY = 7, 8 or 15)
    const int ty = h->adj_ty ? h->adj_ty : (sizeof(T) == 4 ? 7 : 8);
    TileMap tg{(int)h->grd_fused.x, (h->p.ny + ty - 1) / ty, 0, 0};
    tg.run = tg.tx * h->xcd_rows;
    tg.nblk = tg.tx * tg.ty * ngen;
    if (ty == 7)
      hipLaunchKernelGGL((k_adj_gen<T, 7>), dim3(tg.nblk), dim3(kFTX, 8), 0, h->stream, a, gen, tg, rec,
                         h->rec_plane, h->grad);
    else if (ty == 15)
      hipLaunchKernelGGL((k_adj_gen<T, 15>), dim3(tg.nblk), dim3(kFTX, 16), 0, h->stream, a, gen, tg, rec,
                         h->rec_plane, h->grad);
    else
      hipLaunchKernelGGL((k_adj_gen<T, 8>), dim3(tg.nblk), dim3(kFTX, 9), 0, h->stream, a, gen, tg, rec,
                         h->rec_plane, h->grad);
  }
}

// adjoint of step n: A -> B, then A and B swap
template <typename T>
int adjoint_single(iris_handle* h, int n, int nobox, int face_on, const void* rec) {
  ProfScope ps(h, IRIS_K_ADJOINT);
  const StepArgs<T> a = adj_args<T>(h, h->ae, h->ah, h->ae2, h->ah2, h->ad, h->ab, h->ad2, h->ab2);
  adj_sources<T>(h, a, n, nobox, face_on);
  adj_lists<T>(h, a, (const int*)h->chunks, h->nbulk, (const int*)h->chunks + h->nbulk, h->ngen, rec);
  HIP_TRY(h, hipGetLastError());
  std::swap(h->ae, h->ae2);
  std::swap(h->ah, h->ah2);
  std::swap(h->ad, h->ad2);
  std::swap(h->ab, h->ab2);
  return 0;
}

// adjoint of steps n and n-1: the two-step pass on the gradient lists' plain
// chunks (A -> C, exports into B), the other chunks by two single steps
// (A -> B -> C), the sources of step n-1 into B between them
template <typename T>
int adjoint_double(iris_handle* h, int n, int nobox, int face_on, const void* rec_n, const void* rec_n1) {
  using C = cplx<T>;
  ProfScope ps(h, IRIS_K_ADJOINT);
  const ListSet& L = h->al;
  const int* rl = L.rest;
  const int* rg = L.rest + L.nlean;
  const StepArgs<T> a1 = adj_args<T>(h, h->ae, h->ah, h->ae2, h->ah2, h->ad, h->ab, h->ad2, h->ab2);
  adj_sources<T>(h, a1, n, nobox, face_on);
  adj_lists<T>(h, a1, rl, L.nlean, rg, L.ngen, rec_n);
  {
    const StepArgs<T> a2 = adj_args<T>(h, h->ae, h->ah, h->ae3, h->ah3, h->ad, h->ab, h->ad2, h->ab2);
    const TileMap tm{h->tx2, h->ty2, h->tx2 * h->ty2 * L.n2, h->tx2 * h->xcd_rows2};
    Yee2Samples<T> none{};
    if (h->bloch)
      launch_yee2<T, true, false>(h, a2, (C*)h->ae2, (C*)h->ah2, nullptr, (const int*)L.chunks2, tm, L.n2, none, -1,
                                  -1);
    else
      launch_yee2<T, false, false>(h, a2, (C*)h->ae2, (C*)h->ah2, nullptr, (const int*)L.chunks2, tm, L.n2, none, -1,
                                   -1);
  }
  const StepArgs<T> a3 = adj_args<T>(h, h->ae2, h->ah2, h->ae3, h->ah3, h->ad2, h->ab2, h->ad, h->ab);
  adj_sources<T>(h, a3, n - 1, nobox, face_on);
  adj_lists<T>(h, a3, rl, L.nlean, rg, L.ngen, rec_n1);
  HIP_TRY(h, hipGetLastError());
  std::swap(h->ae, h->ae3);
  std::swap(h->ah, h->ah3);
  return 0;
}

// Two-step lists of the gradient runs.  Two-step chunks need every plane
// z0-2 .. z1+2 between the UPML slabs, off the structure and record planes
// hi-1..hf (except in the record lists with two cells a lane, rec: that
// kernel exports H of step n+1 on those planes into the second buffers,
// where the records are taken) and off the TF/SF planes; with face_on (bit 2j+s: face s of pair j
// of box nobox carries weight) also off the planes holding weighted face
// points (the z-faces, or the box's whole z range when an x or y face is
// weighted).  Planes are classified one by one: runs of planes that may be
// two-step outputs become near-equal two-step chunks of up to zc2 planes;
// the planes between take single steps in chunks of up to zc planes that
// carry their own lengths, lean (planes z0-1..z1+1 plain, z1 < nz) or
// general, so the general passes (k_yee_fused<T, true>, k_adj_gen) cover
// only the planes that need them.
template <typename T>
int build_lists(iris_handle* h, ListSet& out, int nobox, int face_on, bool rec = false) {
  const iris_params& p = h->p;
  const int key = nobox * 64 + face_on;
  if (out.key == key && out.gen == h->pass_gen) return 0;
  out.key = key;
  out.gen = h->pass_gen;
  out.on = false;
  out.n2 = out.n2p = out.nlean = out.ngen = 0;
  // record lists with two cells a lane: the record planes may be two-step
  // planes, whose kernel then exports their H of step n+1
  const bool rec_two = rec && sizeof(T) == 4 && h->two_step && h->t2w == 128;
  out.rz0 = rec_two ? p.hi - 1 : 1;
  out.rz1 = rec_two ? p.hf : 0;
  const BoxGeom& g = h->box[nobox - 1].g;
  const bool side = (face_on & 0xf) != 0;
  auto eligible = [&](int z) {
    if (z < 1 || z <= p.pmlzi || z >= p.pmlzf) return false;
    if (!rec_two && z >= p.hi - 1 && z <= p.hf) return false;
    if (p.em_source && (z == p.iz_min - 1 || z == p.iz_min)) return false;
    if (side && z >= g.bzi && z <= g.bzf) return false;
    if (((face_on >> 4) & 1) && z == g.bzi) return false;
    if (((face_on >> 5) & 1) && z == g.bzf) return false;
    return true;
  };
  auto is_bulk_plane = [&](int z) {
    if (z <= p.pmlzi || z >= p.pmlzf) return false;
    if (z >= p.hi && z <= p.hf) return false;
    if (p.em_source && (z == p.iz_min - 1 || z == p.iz_min)) return false;
    return true;
  };
  const bool shape = h->two_step && two_step_fits(p, t2geom<T>(h->t2w, false));
  // output planes of two-step chunks: every plane z-2 .. z+2 eligible
  std::vector<char> two((size_t)p.nz + 2, 0);
  int ndbl = 0;
  for (int z = 3; shape && z + 2 <= p.nz; ++z) {
    bool ok = true;
    for (int q = z - 2; ok && q <= z + 2; ++q) ok = eligible(q);
    two[z] = ok;
  }
  // record lists: the UPML slabs in the UPML variant of the two-step pass
  // where the handle's own lists take it (its output planes as there: no
  // TF/SF plane in z-1 .. z+2, some UPML plane among them), marked 2
  if (rec && h->two_step && h->n2p > 0 && two_step_fits(p, t2geom<T>(h->t2wp, true)))
    for (int z = 2; z <= p.nz - 2; ++z) {
      if (two[z]) continue;
      bool ok = true, pml = false;
      for (int q = z - 1; q <= z + 2; ++q) {
        ok = ok && !(p.em_source && (q == p.iz_min - 1 || q == p.iz_min));
        pml = pml || q <= p.pmlzi || q >= p.pmlzf;
      }
      if (ok && pml) two[z] = 2;
    }
  // runs of fewer than kMinRun2 such planes take single steps
  constexpr int kMinRun2 = 4;
  for (int z = 1; z <= p.nz; ++z) {
    if (!two[z]) continue;
    int e = z;
    while (e + 1 <= p.nz && two[e + 1] == two[z]) ++e;
    if (e - z + 1 < kMinRun2)
      for (int q = z; q <= e; ++q) two[q] = 0;
    else
      ++ndbl;
    z = e;
  }
  if (ndbl == 0) return 0;
  // n near-equal pieces of at most len planes over a .. b
  auto cut = [](int a, int b, int len, auto&& emit) {
    const int n = (b - a + 1 + len - 1) / len, base = (b - a + 1) / n, extra = (b - a + 1) % n;
    for (int k = 0, c0 = a; k < n; ++k) {
      const int l = base + (k < extra ? 1 : 0);
      emit(c0, l);
      c0 += l;
    }
  };
  const int zc2 = h->zc2;
  // single-step chunks carry their length in the chunk word (at most
  // kChunkLenMask planes)
  const int zc = std::min(h->zchunk, kChunkLenMask);
  const int zc2p = std::min(2 * h->zchunk, kChunkLenMask + 1);
  std::vector<int> c2, c2p, rlean, rgen;
  std::vector<std::pair<int, int>> slean, sgen;  // single-step stretches (first, last plane)
  std::vector<unsigned char> zm((size_t)p.nz + 2, 0);
  for (int z = 1; z <= p.nz;) {
    int e = z;
    while (e + 1 <= p.nz && two[e + 1] == two[z]) ++e;
    if (two[z]) {
      // a run of two-step planes: near-equal chunks of up to zc2 planes
      // (zc2p in the UPML variant), the first exporting below, the last above
      std::vector<int> run;
      cut(z, e, two[z] == 2 ? zc2p : zc2, [&](int c0, int l) { run.push_back(c0 | ((l - 1) << kChunkLenShift)); });
      run.front() |= kChunkExportBelow;
      run.back() |= kChunkExportAbove;
      auto& dst = two[z] == 2 ? c2p : c2;
      dst.insert(dst.end(), run.begin(), run.end());
      for (int q = z; q <= e; ++q) zm[q] = 1;
    } else {
      // single-step planes: lean where the planes a chunk reads or writes
      // are all bulk (and it is not the top plane), general elsewhere;
      // lean stretches shorter than kMinRun2 planes go with the general
      auto lean_ok = [&](int q) { return q < p.nz && is_bulk_plane(q - 1) && is_bulk_plane(q) && is_bulk_plane(q + 1); };
      std::vector<char> ln((size_t)(e - z + 1));
      for (int q = z; q <= e; ++q) ln[q - z] = lean_ok(q);
      for (int q = z; q <= e;) {
        int f = q;
        while (f + 1 <= e && ln[f + 1 - z] == ln[q - z]) ++f;
        if (ln[q - z] && f - q + 1 < kMinRun2)
          for (int r = q; r <= f; ++r) ln[r - z] = 0;
        q = f + 1;
      }
      for (int q = z; q <= e;) {
        int f = q;
        while (f + 1 <= e && ln[f + 1 - z] == ln[q - z]) ++f;
        (ln[q - z] ? slean : sgen).push_back({q, f});
        q = f + 1;
      }
    }
    z = e + 1;
  }
  // single-step stretches in chunks of up to zc planes.  The adjoint's
  // general chunks (k_adj_gen) take shorter ones (down to 4 planes) where
  // zc-plane chunks would give fewer than 8 workgroups per CU: they cover a
  // few thin slabs (UPML, structure), which in zc-plane chunks leave most
  // CUs idle (on 256 x 256 x 1024, 84 planes in 5 chunks of 128 tiles: 460
  // against 369 us a step in 6-plane chunks).  The forward's general pass
  // measured no faster in shorter chunks (177 against 188 us there).
  auto emit = [&](const std::vector<std::pair<int, int>>& segs, std::vector<int>& dst, bool fill) {
    const int64_t tiles = (int64_t)h->grd_fused.x * h->grd_fused.y, want = 8LL * h->ncu;
    auto wgs = [&](int l) {
      int64_t n = 0;
      for (const auto& sg : segs) n += (sg.second - sg.first + l) / l;
      return n * tiles;
    };
    int L = zc;
    while (fill && L > 4 && wgs(L) < want) --L;
    for (const auto& sg : segs) cut(sg.first, sg.second, L, [&](int c0, int l) { dst.push_back(c0 | (l << kChunkLenShift)); });
  };
  emit(slean, rlean, false);
  emit(sgen, rgen, !rec);
  std::vector<int> rest(rlean);
  rest.insert(rest.end(), rgen.begin(), rgen.end());
  int rc;
  if (!out.chunks2) {
    // sized for the most chunks any face mask can give (one a plane)
    const size_t cap = (size_t)p.nz + 1;
    if ((rc = adj_alloc(h, (void**)&out.chunks2, sizeof(int) * cap)) ||
        (rc = adj_alloc(h, (void**)&out.chunks2p, sizeof(int) * cap)) ||
        (rc = adj_alloc(h, (void**)&out.rest, sizeof(int) * cap)) || (rc = adj_alloc(h, (void**)&out.zmask, zm.size())))
      return rc;
  }
  if (!c2.empty())
    HIP_TRY(h, hipMemcpyAsync(out.chunks2, c2.data(), sizeof(int) * c2.size(), hipMemcpyHostToDevice, h->stream));
  if (!c2p.empty())
    HIP_TRY(h, hipMemcpyAsync(out.chunks2p, c2p.data(), sizeof(int) * c2p.size(), hipMemcpyHostToDevice, h->stream));
  if (!rest.empty())
    HIP_TRY(h, hipMemcpyAsync(out.rest, rest.data(), sizeof(int) * rest.size(), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipMemcpyAsync(out.zmask, zm.data(), zm.size(), hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  out.n2 = (int)c2.size();
  out.n2p = (int)c2p.size();
  out.nlean = (int)rlean.size();
  out.ngen = (int)rgen.size();
  out.on = true;
  return 0;
}

// Steps it_first..it_last of the recorded tail (all after step tail_t0): H of
// every step goes to its record slot; two steps per pass on the record
// lists (the record planes are in their single-step chunks, or in two-step
// chunks of two cells a lane that export their H of the first step, so that
// H is whole there in the second buffer), the DFT as in run_steps.
template <typename T>
int run_tail(iris_handle* h, int it_first, int it_last) {
  const iris_params& p = h->p;
  int it = it_first;
  auto rec = [&](const void* f, int step) -> int {
    const int r = step - h->tail_t0;
    if (r < 0 || r >= h->tail_n || h->tail_ok[r]) return 0;
    h->tail_ok[r] = 1;
    return record_h(h, f, h->tail_rec, r);
  };
  int rc = rec(h->fh, it - 1);
  if (rc) return rc;
  if (h->rl.key >= 0 && (rc = build_lists<T>(h, h->rl, 1, 0, true))) return rc;  // a rebuilt layout since
  const PassLists L = h->rl.view();
  while (h->rl.on && it_last - it + 1 >= 3) {
    if ((rc = launch_double<T>(h, it, L, true))) return rc;
    if ((rc = rec(h->gh, it)) || (rc = rec(h->fh, it + 1))) return rc;
    it += 2;
  }
  for (; it <= it_last; ++it) {
    if ((rc = launch_step<T>(h, it))) return rc;
    if (it >= p.to_fourier && (it % p.sample_fourier) == 0 && (rc = dft_sample<T>(h, it + 1 - p.to_fourier)))
      return rc;
    h->it = it;
    if ((rc = rec(h->fh, it))) return rc;
  }
  return 0;
}

template <typename T>
int adjoint_impl(iris_handle* h, int nobox, const double* weights, double* grad_out) {
  const iris_params& p = h->p;
  const int nt = p.nt, seg = h->ck_seg;
  const int nhead = (int)h->ck.size();
  const size_t fb = field_bytes(h), abytes = aux_bytes(h);
  const bool fused = h->fused && h->adj_fused;
  int rc;
  if ((rc = flush(h))) return rc;
  // buffers
  if (!h->ae) {
    const size_t gbytes = sizeof(double2) * (size_t)p.nx * p.ny * (p.hf - p.hi + 1);
    if ((rc = adj_alloc(h, &h->ae, fb)) || (rc = adj_alloc(h, &h->ah, fb)) || (rc = adj_alloc(h, &h->atmp, fb)) ||
        (rc = adj_alloc(h, &h->ad, abytes)) || (rc = adj_alloc(h, &h->ab, abytes)) ||
        (rc = adj_alloc(h, (void**)&h->grad, gbytes)) ||
        (rc = adj_alloc(h, (void**)&h->wdev, sizeof(double) * 6 * p.nof)) ||
        (rc = adj_alloc(h, &h->ck_final, state_bytes(h))))
      return rc;
  }
  if (fused && !h->ae2) {
    if ((rc = adj_alloc(h, &h->ae2, fb)) || (rc = adj_alloc(h, &h->ah2, fb)) || (rc = adj_alloc(h, &h->ad2, abytes)) ||
        (rc = adj_alloc(h, &h->ab2, abytes)))
      return rc;
  }
  if (nhead > 0 && h->rec_cap < seg) {
    if ((rc = adj_alloc(h, &h->hrec, rec_step_bytes(h) * seg))) return rc;
    h->rec_cap = seg;
  }
  const DeviceBox& B = h->box[nobox - 1];
  const int64_t gneed = std::max<int64_t>(B.g.ntot, 1) * p.nof;
  if (gneed > h->adj_g_cap) {
    if ((rc = adj_alloc(h, (void**)&h->adj_g, sizeof(double2) * (size_t)gneed))) return rc;
    h->adj_g_cap = gneed;
  }
  // dJ/d(emg) of the weighted faces
  int face_on = 0;
  for (int j = 0; j < 3; ++j)
    for (int s = 0; s < 2; ++s) {
      bool any = false;
      for (int iw = 0; iw < p.nof; ++iw) any = any || weights[iw * 6 + 2 * j + s] != 0.0;
      if (any && B.g.npts[j]) face_on |= 1 << (2 * j + s);
    }
  HIP_TRY(h, hipMemcpyAsync(h->wdev, weights, sizeof(double) * 6 * p.nof, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipMemsetAsync(h->adj_g, 0, sizeof(double2) * (size_t)gneed, h->stream));
  const double sflux = (p.eps0 * p.q) / (2.0 * p.dt);
  for (int j = 0; j < 3; ++j)
    if (B.g.npts[j])
      hipLaunchKernelGGL(k_adj_g, dim3(blocks_for(B.g.npts[j])), dim3(256), 0, h->stream, B.g, j,
                         (const double2*)B.emg, p.nof, (const double*)h->wdev, sflux, h->adj_g);
  HIP_TRY(h, hipGetLastError());
  if (fused && (rc = build_lists<T>(h, h->al, nobox, face_on))) return rc;
  if (fused && h->al.on && !h->ae3) {
    if ((rc = adj_alloc(h, &h->ae3, fb)) || (rc = adj_alloc(h, &h->ah3, fb))) return rc;
  }
  // zero adjoint state (ghosts included) and gradient; keep the state at nt
  for (void* f : {h->ae, h->ah, h->atmp}) HIP_TRY(h, hipMemsetAsync(f, 0, fb, h->stream));
  for (void* f : {h->ad, h->ab}) HIP_TRY(h, hipMemsetAsync(f, 0, abytes, h->stream));
  if (fused) {
    for (void* f : {h->ae2, h->ah2}) HIP_TRY(h, hipMemsetAsync(f, 0, fb, h->stream));
    for (void* f : {h->ad2, h->ab2}) HIP_TRY(h, hipMemsetAsync(f, 0, abytes, h->stream));
    if (h->ae3)
      for (void* f : {h->ae3, h->ah3}) HIP_TRY(h, hipMemsetAsync(f, 0, fb, h->stream));
  }
  const size_t gbytes = sizeof(double2) * (size_t)p.nx * p.ny * (p.hf - p.hi + 1);
  HIP_TRY(h, hipMemsetAsync(h->grad, 0, gbytes, h->stream));
  if ((rc = state_copy(h, h->ck_final, true))) return rc;

  // adjoint of steps lo+1..hi_, H of step m-1 in slot m-1-lo of `base`
  auto sweep = [&](int lo, int hi_, const void* base) -> int {
    for (int m = hi_; m > lo;) {
      int r;
      if (!fused) {
        r = adjoint_step<T>(h, m, nobox, face_on, rec_at(h, base, m - 1 - lo));
        m -= 1;
      } else if (h->al.on && m - lo >= 2) {
        r = adjoint_double<T>(h, m, nobox, face_on, rec_at(h, base, m - 1 - lo), rec_at(h, base, m - 2 - lo));
        m -= 2;
      } else {
        r = adjoint_single<T>(h, m, nobox, face_on, rec_at(h, base, m - 1 - lo));
        m -= 1;
      }
      if (r) return r;
    }
    return 0;
  };
  // the recorded tail needs no recomputation
  if (h->tail_n && (rc = sweep(h->tail_t0, nt, h->tail_rec))) return rc;
  if (h->rl.key >= 0 && (rc = build_lists<T>(h, h->rl, 1, 0, true))) return rc;  // a rebuilt layout since
  const PassLists RL = h->rl.view();
  for (int k = nhead - 1; k >= 0; --k) {
    const int s0 = k * seg, s1 = std::min(h->tail_t0, s0 + seg);
    // forward again from the snapshot, recording H of steps s0..s1-1 (two
    // steps per pass on the record lists: H of the first step is then whole
    // on the record planes in the second buffer)
    if ((rc = state_copy(h, h->ck[k], false))) return rc;
    if ((rc = record_h(h, h->fh, h->hrec, 0))) return rc;
    int n = s0 + 1;
    if (fused && h->rl.on)
      for (; n + 1 < s1; n += 2) {
        if ((rc = launch_double<T>(h, n, RL, false))) return rc;
        if ((rc = record_h(h, h->gh, h->hrec, n - s0)) || (rc = record_h(h, h->fh, h->hrec, n + 1 - s0))) return rc;
      }
    for (; n < s1; ++n) {
      if ((rc = launch_step<T>(h, n))) return rc;
      if ((rc = record_h(h, h->fh, h->hrec, n - s0))) return rc;
    }
    if ((rc = sweep(s0, s1, h->hrec))) return rc;
  }
  if ((rc = state_copy(h, h->ck_final, false))) return rc;
  h->it = nt;
  HIP_TRY(h, hipMemcpyAsync(grad_out, h->grad, gbytes, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

// ---- layout tuning ---------------------------------------------------------------

// Microseconds per step of the current layout: `steps` steps (pairs where
// the two-step pass applies) after one warm-up pair, without DFT samples,
// timed with events on the handle's stream.  Changes the field state.
template <typename T>
int trial_steps(iris_handle* h, int steps, double* us) {
  PassLists L;
  L.chunks2 = h->chunks2;
  L.n2 = h->n2;
  L.rest = h->chunks_rest;
  L.nlean = h->nrest_lean;
  L.ngen = h->nrest_gen;
  L.zmask = h->zmask;
  L.chunks2p = h->chunks2p;
  L.n2p = h->n2p;
  L.groups = h->rest_groups.data();
  L.ngroups = (int)h->rest_groups.size();
  auto run = [&](int n) -> int {
    for (int it = 1; it <= n;) {
      int rc;
      if (h->two_step && n - it + 1 >= 2) {
        if ((rc = launch_double<T>(h, it, L, false))) return rc;
        it += 2;
      } else {
        if ((rc = launch_step<T>(h, it))) return rc;
        ++it;
      }
    }
    return 0;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(h, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(h, IRIS_EDEVICE, "hipEventCreate failed");
  }
  int rc = run(2);
  if (!rc) {
    (void)hipEventRecord(e0, h->stream);
    rc = run(steps);
    (void)hipEventRecord(e1, h->stream);
  }
  float ms = 0.f;
  if (!rc && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
    *us = 1000.0 * ms / steps;
  else if (!rc)
    rc = fail(h, IRIS_EDEVICE, "timing the layout failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Picks the fastest layout by timing candidates one parameter at a time
// (coordinate descent from the current layout; a change is kept when it is
// at least 2% faster).  Only layouts that give the same results bit for bit
// are compared: with a Bloch phase the choice between the two-step pass and
// single steps (and the UPML variant) stays as it is.  The field state is
// saved before and restored after; the handle stays at step 0.
template <typename T>
int autotune_impl(iris_handle* h, int steps, int64_t info[9]) {
  const size_t bytes = state_bytes(h);
  void* snap = nullptr;
  if (hipMalloc(&snap, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, IRIS_ENOMEM, "iris_autotune: no device memory for a copy of the state (%zu bytes)", bytes);
  }
  const bool prof = h->prof;
  h->prof = false;
  int rc = state_copy(h, snap, true);
  int tried = 0;
  PassConfig best = h->pass;
  double tbest = 0.0;
  auto measure = [&](const PassConfig& c, double* us) -> int {
    int r = build_passes<T>(h, c);
    if (!r) r = trial_steps<T>(h, steps, us);
    ++tried;
    return r;
  };
  if (!rc) rc = measure(best, &tbest);
  auto consider = [&](PassConfig c) {
    if (rc) return;
    double t = 0.0;
    rc = measure(c, &t);
    if (!rc && t < 0.98 * tbest) {
      best = c;
      tbest = t;
    }
  };
  const iris_params& p = h->p;
  const int zc0 = h->zchunk;
  if (!rc && !h->bloch) {
    for (int v : {0, 1}) {
      PassConfig c = best;
      c.two = v;
      if (c.two != best.two && !(best.two < 0 && (v == 1) == h->two_step)) consider(c);
    }
    if (!rc) rc = build_passes<T>(h, best);
    if (!rc && h->two_step)
      for (int v : {0, 1}) {
        PassConfig c = best;
        c.upml = v;
        if (c.upml != best.upml && !(best.upml < 0 && (v == 1) == (h->upml2p > 0))) consider(c);
      }
  }
  if (!rc) rc = build_passes<T>(h, best);
  const bool two = h->two_step;
  if (!rc && two) {  // tile width and output stores of the two-step pass
    // with a Bloch phase the tile width stays: the lanes that stand for
    // cells across the x period take their neighbours from other lanes or
    // from the ghost planes depending on the tiles, the same values up to
    // the rounding of the phase factors
    const int w0 = h->t2w;
    for (int v : {64, 32, 128}) {
      PassConfig c = best;
      c.w2 = v;
      if (!h->bloch && v != w0 && (v != 128 || (sizeof(T) == 4 && h->p.nx % 2 == 0))) consider(c);
    }
    if (!rc) rc = build_passes<T>(h, best);
    const int wp0 = h->t2wp;
    if (!rc && h->upml2p > 0 && !h->bloch)
      for (int v : {64, 32}) {
        PassConfig c = best;
        c.w2p = v;
        if (v != wp0) consider(c);
      }
    if (!rc) rc = build_passes<T>(h, best);
    const int n0 = h->nt2;
    if (!rc) {
      PassConfig c = best;
      c.nt2 = 1 - n0;
      consider(c);
    }
    if (!rc) rc = build_passes<T>(h, best);
  }
  for (int f : {1, 2}) {  // single-step chunks: shorter, longer
    PassConfig c = best;
    c.zc = f == 1 ? zc0 / 2 : std::min(2 * zc0, p.nz);
    if (c.zc >= 1 && c.zc != zc0) consider(c);
  }
  for (int v : {2, 4}) {
    PassConfig c = best;
    c.xcd_rows = v;
    if (v != best.xcd_rows) consider(c);
  }
  if (two) {
    if (!rc) rc = build_passes<T>(h, best);
    const int zc2 = h->zc2;
    for (int v : {zc2 / 2, 2 * zc2}) {
      PassConfig c = best;
      c.zc2 = std::min(v, kChunkLenMask + 1);
      if (c.zc2 >= 2 && c.zc2 != zc2) consider(c);
    }
    const int xr2 = h->xcd_rows2;
    for (int v : {0, 2, 4, 8}) {
      PassConfig c = best;
      c.xcd_rows2 = v;
      if (v != xr2) consider(c);
    }
  }
  const int rc_run = rc;
  rc = build_passes<T>(h, best);
  if (!rc) rc = state_copy(h, snap, false);
  // the other buffers as a new handle has them (the previous step's fields
  // read at step 0 are zero)
  for (void* q : {h->ge, h->gh})
    if (!rc && q && hipMemsetAsync(q, 0, field_bytes(h), h->stream) != hipSuccess)
      rc = fail(h, IRIS_EDEVICE, "hipMemsetAsync failed");
  for (void* q : {h->dn, h->bn})
    if (!rc && q && hipMemsetAsync(q, 0, aux_bytes(h), h->stream) != hipSuccess)
      rc = fail(h, IRIS_EDEVICE, "hipMemsetAsync failed");
  if (!rc) HIP_TRY(h, hipStreamSynchronize(h->stream));
  (void)hipFree(snap);
  h->prof = prof;
  h->it = 0;
  if (rc_run) return rc_run;
  if (rc) return rc;
  info[0] = h->two_step ? 1 : 0;
  info[1] = h->two_step ? h->upml2p : 0;
  info[2] = h->zchunk;
  info[3] = h->two_step ? h->zc2 : 0;
  info[4] = h->xcd_rows;
  info[5] = h->xcd_rows2;
  info[6] = tried;
  info[7] = (int64_t)(tbest * 1000.0 + 0.5);
  info[8] = h->two_step ? h->t2w : 0;
  return 0;
}

// ---- batches ------------------------------------------------------------------

int bfail(iris_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf; else g_create_error = buf;
  return code;
}

#define BHIP_TRY(b, call)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return bfail((b), IRIS_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

int balloc(iris_batch* b, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return bfail(b, IRIS_ENOMEM, "hipMalloc(%zu bytes) for the batch failed", bytes);
  }
  b->allocs.push_back(*p);
  return 0;
}

// argument tables of the current buffers of every instance (parity 0:
// current -> other; 1: the reverse), uploaded before a run of steps
template <typename T>
int batch_tables(iris_batch* b) {
  const int n = (int)b->one.size();
  if (n == 0) return 0;
  b->host_inst.assign(sizeof(BatchInst<T>) * n, 0);
  BatchInst<T>* I = (BatchInst<T>*)b->host_inst.data();
  std::vector<BatchSamp<T>> E;
  for (int i = 0; i < n; ++i) {
    iris_handle* h = b->one[i];
    StepArgs<T> a0 = step_args_on<T>(h, h->fe, h->fh, h->ge, h->gh, h->d, h->b, h->dn, h->bn);
    StepArgs<T> a1 = step_args_on<T>(h, h->ge, h->gh, h->fe, h->fh, h->dn, h->bn, h->d, h->b);
    a0.zchunk = a1.zchunk = b->zc;
    I[i].a[0] = a0;
    I[i].a[1] = a1;
    I[i].tp = (const TfsfConst*)h->tfc;
    I[i].chunks[0] = b->lean[i];
    I[i].chunks[1] = b->gen[i];
    const int tx = (int)h->grd_fused.x, ty = (int)h->grd_fused.y;
    I[i].tm[0] = TileMap{tx, ty, tx * ty * b->nlean[i], tx * h->xcd_rows};
    I[i].tm[1] = TileMap{tx, ty, tx * ty * b->ngen[i], tx * h->xcd_rows};
    for (int bx = 0; bx < h->p.nbox; ++bx) {
      if (h->box[bx].g.ntot == 0) continue;
      BatchSamp<T> e{};
      e.a[0] = a0;
      e.a[1] = a1;
      e.g = h->box[bx].g;
      e.samp = (cplx<T>*)h->box[bx].samp;
      E.push_back(e);
    }
  }
  b->host_ent.assign((const char*)E.data(), (const char*)(E.data() + E.size()));
  BHIP_TRY(b, hipMemcpy(b->inst, b->host_inst.data(), b->host_inst.size(), hipMemcpyHostToDevice));
  if (!E.empty()) BHIP_TRY(b, hipMemcpy(b->ent, b->host_ent.data(), b->host_ent.size(), hipMemcpyHostToDevice));
  return 0;
}

// argument tables of the two-step handles for their current buffers: the
// pair of steps from the current buffers (pair parity 0) or from the third
// ones (parity 1); step n+1 always goes to the second buffers
template <typename T>
int batch_tables2(iris_batch* b) {
  using C = cplx<T>;
  if (b->own.empty()) return 0;
  for (int bl = 0; bl < 16; ++bl) {
    for (int pm = 0; pm < 2; ++pm) {
      const auto& es = b->y2e[bl][pm];
      const int n = (int)es.size();
      if (!n) continue;
      std::vector<Batch2Inst<T>> I((size_t)2 * n);
      for (int pp = 0; pp < 2; ++pp)
        for (int i = 0; i < n; ++i) {
          iris_handle* h = es[i].first;
          const int part = es[i].second;
          Batch2Inst<T>& B = I[(size_t)pp * n + i];
          void *A0 = pp ? h->ke : h->fe, *A1 = pp ? h->kh : h->fh, *C0 = pp ? h->fe : h->ke, *C1 = pp ? h->fh : h->kh;
          void *AD = pp ? h->kd : h->d, *AB = pp ? h->kb : h->b, *CD = pp ? h->d : h->kd, *CB = pp ? h->b : h->kb;
          B.a = pm ? step_args_on<T>(h, A0, A1, C0, C1, AD, AB, CD, CB)
                   : step_args_on<T>(h, A0, A1, C0, C1, AD, AB, h->dn, h->bn);
          B.eb = (C*)h->ge;
          B.hb = (C*)h->gh;
          B.db = pm ? (C*)h->dn : nullptr;
          B.chunks = pm ? h->chunks2p : h->chunks2;
          const int nch = pm ? h->n2p : h->n2;
          if (part) {
            B.tm = TileMap{h->tx2n, h->ty2n, h->tx2n * h->ty2n * nch, h->tx2n * h->xcd_rows2, h->x0n};
            B.ox = h->p2oxn;
            B.sl = h->p2sln;
            B.ns = h->p2nsn;
          } else {
            const int tx = pm ? h->tx2p : h->tx2, ty = pm ? h->ty2p : h->ty2;
            B.tm = TileMap{tx, ty, tx * ty * nch, tx * h->xcd_rows2, 0};
            B.ox = h->p2ox;
            B.sl = h->p2sl;
            B.ns = h->p2ns;
          }
          B.smp = Yee2Samples<T>{};
          B.smp.nbox = h->p.nbox;
          for (int bx = 0; bx < h->p.nbox; ++bx) {
            B.smp.g[bx] = h->box[bx].g;
            B.smp.samp[bx] = h->box[bx].g.ntot ? (C*)h->box[bx].samp : nullptr;
          }
        }
      BHIP_TRY(b, hipMemcpy(b->y2[bl][pm], I.data(), sizeof(Batch2Inst<T>) * I.size(), hipMemcpyHostToDevice));
    }
  }
  // single steps: entry (handle, length group); a[0] = step n+1, a[1] = step n+2
  std::vector<BatchInst<T>> R((size_t)2 * b->nr);
  std::vector<BatchSamp<T>> E((size_t)2 * b->nent2);
  for (int pp = 0; pp < 2; ++pp) {
    int r = 0, e = 0;
    for (iris_handle* h : b->own) {
      void *A0 = pp ? h->ke : h->fe, *A1 = pp ? h->kh : h->fh, *C0 = pp ? h->fe : h->ke, *C1 = pp ? h->fh : h->kh;
      void *AD = pp ? h->kd : h->d, *AB = pp ? h->kb : h->b, *CD = pp ? h->d : h->kd, *CB = pp ? h->b : h->kb;
      const StepArgs<T> s1 = step_args_on<T>(h, A0, A1, h->ge, h->gh, AD, AB, h->dn, h->bn);
      const StepArgs<T> s2 = step_args_on<T>(h, h->ge, h->gh, C0, C1, h->dn, h->bn, CD, CB);
      const int tx = (int)h->grd_fused.x, ty = (int)h->grd_fused.y;
      for (const RestGroup& G : h->rest_groups) {
        BatchInst<T>& I = R[(size_t)pp * b->nr + r++];
        I.a[0] = s1;
        I.a[1] = s2;
        I.a[0].zchunk = I.a[1].zchunk = G.len;
        I.tp = (const TfsfConst*)h->tfc;
        I.chunks[0] = h->chunks_rest + G.off;
        I.chunks[1] = h->chunks_rest + G.off + G.nlean;
        I.tm[0] = TileMap{tx, ty, tx * ty * G.nlean, tx * h->xcd_rows};
        I.tm[1] = TileMap{tx, ty, tx * ty * G.ngen, tx * h->xcd_rows};
      }
      // samples: pp 0 = step n+1 from the second buffers outside the
      // two-step planes; pp 1 = step n+2, a[par] by the pair's parity
      for (int bx = 0; bx < h->p.nbox; ++bx) {
        if (h->box[bx].g.ntot == 0) continue;
        BatchSamp<T>& S = E[(size_t)pp * b->nent2 + e++];
        S.g = h->box[bx].g;
        S.samp = (C*)h->box[bx].samp;
        if (pp == 0) {
          S.a[0] = S.a[1] = step_args_on<T>(h, h->ge, h->gh, h->ge, h->gh, h->dn, h->bn, h->dn, h->bn);
          S.zmask = h->zmask;
          S.xmask = nullptr;
        } else {
          S.a[0] = step_args_on<T>(h, h->ke, h->kh, h->ke, h->kh, h->kd, h->kb, h->kd, h->kb);
          S.a[1] = step_args_on<T>(h, h->fe, h->fh, h->fe, h->fh, h->d, h->b, h->d, h->b);
          S.zmask = nullptr;
          S.xmask = h->zmask;  // with xm: the x faces of the two-step planes are stored by their kernels
        }
      }
    }
  }
  if (b->nr) BHIP_TRY(b, hipMemcpy(b->rinst, R.data(), sizeof(BatchInst<T>) * R.size(), hipMemcpyHostToDevice));
  if (b->nent2) BHIP_TRY(b, hipMemcpy(b->ent2, E.data(), sizeof(BatchSamp<T>) * E.size(), hipMemcpyHostToDevice));
  return 0;
}

// workgroup prefixes and table memory of the batched two-step handles
template <typename T>
int batch_create2(iris_batch* b) {
  int rc;
  for (iris_handle* h : b->own) {
    const int bl = (h->bloch ? 1 : 0) + (h->t2w == 32 ? 2 : 0) + (h->t2wp == 32 ? 4 : 0) + (h->t2w == 128 ? 8 : 0);
    b->own2[bl].push_back(h);
    if (h->n2) {
      if (h->t2w != 128 || h->tx2) b->y2e[bl][0].push_back({h, 0});
      if (h->t2w == 128 && h->tx2n) b->y2e[bl][0].push_back({h, 1});
    }
    if (h->n2p) b->y2e[bl][1].push_back({h, 0});
  }
  for (int bl = 0; bl < 16; ++bl) {
    for (int pm = 0; pm < 2; ++pm) {
      const auto& es = b->y2e[bl][pm];
      if (es.empty()) continue;
      std::vector<int> wg = {0};
      for (const auto& e : es) {
        iris_handle* h = e.first;
        const int64_t nb = e.second ? (int64_t)h->tx2n * h->ty2n * h->n2
                                    : (int64_t)(pm ? h->tx2p : h->tx2) * (pm ? h->ty2p : h->ty2) * (pm ? h->n2p : h->n2);
        wg.push_back(wg.back() + (int)((nb + kNumXcd - 1) / kNumXcd * kNumXcd));
      }
      if ((rc = balloc(b, (void**)&b->y2wg[bl][pm], sizeof(int) * wg.size())) ||
          (rc = balloc(b, &b->y2[bl][pm], sizeof(Batch2Inst<T>) * 2 * es.size())))
        return rc;
      BHIP_TRY(b, hipMemcpy(b->y2wg[bl][pm], wg.data(), sizeof(int) * wg.size(), hipMemcpyHostToDevice));
      b->y2nwg[bl][pm] = wg.back();
    }
  }
  std::vector<int> wg[2] = {{0}, {0}};
  std::vector<int64_t> start = {0};
  for (iris_handle* h : b->own) {
    const int t = (int)(h->grd_fused.x * h->grd_fused.y);
    for (const RestGroup& G : h->rest_groups) {
      wg[0].push_back(wg[0].back() + t * G.nlean);
      wg[1].push_back(wg[1].back() + t * G.ngen);
      b->nr++;
    }
    for (int bx = 0; bx < h->p.nbox; ++bx)
      if (h->box[bx].g.ntot) start.push_back(start.back() + (h->box[bx].g.ntot + 255) / 256 * 256);
  }
  for (int k = 0; k < 2; ++k) {
    if ((rc = balloc(b, (void**)&b->rwg[k], sizeof(int) * wg[k].size()))) return rc;
    BHIP_TRY(b, hipMemcpy(b->rwg[k], wg[k].data(), sizeof(int) * wg[k].size(), hipMemcpyHostToDevice));
    b->rnwg[k] = wg[k].back();
  }
  b->nent2 = (int)start.size() - 1;
  b->npts2 = start.back();
  if ((rc = balloc(b, (void**)&b->ent2_start, sizeof(int64_t) * start.size())) ||
      (rc = balloc(b, &b->rinst, sizeof(BatchInst<T>) * 2 * std::max(b->nr, 1))) ||
      (rc = balloc(b, &b->ent2, sizeof(BatchSamp<T>) * 2 * std::max(b->nent2, 1))))
    return rc;
  BHIP_TRY(b, hipMemcpy(b->ent2_start, start.data(), sizeof(int64_t) * start.size(), hipMemcpyHostToDevice));
  return 0;
}

// one DFT sample of every two-step handle of the batch (table half `half`,
// state a[par]), then each handle's ring bookkeeping as dft_sample does it
template <typename T>
int batch_sample2(iris_batch* b, int half, int par, int n, int xm) {
  iris_handle* h0 = b->own[0];
  if (b->nent2)
    hipLaunchKernelGGL(k_dft_sample_batch<T>, dim3((unsigned)((b->npts2 + 255) / 256)), dim3(256), 0, b->stream,
                       (const BatchSamp<T>*)b->ent2 + (size_t)half * b->nent2, (const int64_t*)b->ent2_start,
                       b->nent2, par, h0->nsamp, xm);
  for (iris_handle* h : b->own) {
    if (h->nsamp == 0) h->samp_n0 = n;
    h->nsamp++;
    if (h->nsamp == dft_slots<T>()) {
      const int rc = dft_flush<T>(h);
      if (rc) {
        b->err = h->err;
        return rc;
      }
    }
  }
  return 0;
}

template <typename T, bool PH, bool PML>
void launch_yee2_batch(const iris_batch* b, int W, unsigned nblk, const Batch2Inst<T>* I, const int* wg, int n,
                       int slot, int slot2) {
  if constexpr (sizeof(T) == 4 && !PML) {
    if (W == 128) {
      hipLaunchKernelGGL((k_yee2p_batch<PH>), dim3(nblk), dim3(64, kP2R), 0, b->stream, I, wg, n, slot, slot2);
      return;
    }
  }
  if (W == 32)
    hipLaunchKernelGGL((k_yee2_batch<T, PH, PML, 32>), dim3(nblk), dim3(kT2X, Yee2Shape<T, PML, 32>::R), 0,
                       b->stream, I, wg, n, slot, slot2);
  else
    hipLaunchKernelGGL((k_yee2_batch<T, PH, PML, 64>), dim3(nblk), dim3(kT2X, Yee2Shape<T, PML, 64>::R), 0,
                       b->stream, I, wg, n, slot, slot2);
}

// steps it_first..it_last of the two-step handles: pairs of steps while at
// least three remain (as run_steps takes them alone), then single steps
template <typename T>
int batch_run2(iris_batch* b, int it_first, int it_last) {
  if (b->own.empty()) return 0;
  const iris_params& p = b->own[0]->p;
  int it = it_first, pp = 0, rc = 0;
  auto rest = [&](int k, int it_step) {
    for (int c = 0; c < 2; ++c) {
      if (!b->rnwg[c]) continue;
      const BatchInst<T>* I = (const BatchInst<T>*)b->rinst + (size_t)pp * b->nr;
      if (c == 0)
        hipLaunchKernelGGL((k_yee_fused_batch<T, false>), dim3(b->rnwg[0]), dim3(kFTX, kFTY + 1), 0, b->stream, I,
                           (const int*)b->rwg[0], b->nr, k, it_step);
      else
        hipLaunchKernelGGL((k_yee_fused_batch<T, true>), dim3(b->rnwg[1]), dim3(kFTX, kFTY + 1), 0, b->stream, I,
                           (const int*)b->rwg[1], b->nr, k, it_step);
    }
  };
  // one pair of steps from `it` (updates the pair parity)
  auto pair = [&](int it) -> int {
    const bool s1 = it >= p.to_fourier && (it % p.sample_fourier) == 0;
    const bool s2 = it + 1 >= p.to_fourier && ((it + 1) % p.sample_fourier) == 0;
    const int slot = s1 ? b->own[0]->nsamp : -1;
    // the x faces of step n+2 by the two-step kernels (as launch_double)
    const int ns2 = b->own[0]->nsamp + (s1 ? 1 : 0);
    const int slot2 = (s2 && ns2 < dft_slots<T>()) ? ns2 : -1;
    int rc;
    rest(0, it);
    for (int bl = 0; bl < 16; ++bl) {
      for (int pm = 0; pm < 2; ++pm) {
        if (!b->y2nwg[bl][pm]) continue;
        const int n = (int)b->y2e[bl][pm].size();
        const Batch2Inst<T>* I = (const Batch2Inst<T>*)b->y2[bl][pm] + (size_t)pp * n;
        const int* w = b->y2wg[bl][pm];
        const int W = (!pm && (bl & 8)) ? 128 : ((bl & (pm ? 4 : 2)) ? 32 : 64);
        const unsigned g = (unsigned)b->y2nwg[bl][pm];
        if (bl & 1) {
          if (pm) launch_yee2_batch<T, true, true>(b, W, g, I, w, n, slot, slot2);
          else launch_yee2_batch<T, true, false>(b, W, g, I, w, n, slot, slot2);
        } else {
          if (pm) launch_yee2_batch<T, false, true>(b, W, g, I, w, n, slot, slot2);
          else launch_yee2_batch<T, false, false>(b, W, g, I, w, n, slot, slot2);
        }
      }
    }
    rest(1, it + 1);
    if (hipGetLastError() != hipSuccess) return bfail(b, IRIS_EDEVICE, "batched two-step launch failed");
    if (s1 && (rc = batch_sample2<T>(b, 0, 0, it + 1 - p.to_fourier, 0))) return rc;
    if (s2 && (rc = batch_sample2<T>(b, 1, pp, it + 2 - p.to_fourier, slot2 >= 0 ? 1 : 0))) return rc;
    pp ^= 1;
    return 0;
  };
  // periods of pairs as graph replays where they fit (see iris_batch)
  const int P = dft_slots<T>() * p.sample_fourier;
  bool graph_ok = b->own[0]->graph_on;
  for (iris_handle* h : b->own) graph_ok = graph_ok && !h->prof;
  bool shifted = false;
  auto set_shift = [&](int s) -> int {
    for (iris_handle* h : b->own) hipLaunchKernelGGL(k_set_shift, dim3(1), dim3(64), 0, b->stream, h->tfc, s);
    if (hipGetLastError() != hipSuccess) return bfail(b, IRIS_EDEVICE, "step shift launch failed");
    shifted = s != 0;
    return 0;
  };
  while (it_last - it + 1 >= 3) {
    if (graph_ok && pp == 0 && b->own[0]->nsamp == 0 && it >= p.to_fourier && it_last - it + 1 >= P + 1) {
      if (b->gexec && (it - b->g_it0) % p.sample_fourier == 0) {
        if ((rc = set_shift(it - b->g_it0))) return rc;
        if (hipGraphLaunch(b->gexec, b->stream) != hipSuccess) {
          (void)set_shift(0);
          return bfail(b, IRIS_EDEVICE, "hipGraphLaunch failed");
        }
        b->g_replays++;
        for (iris_handle* h : b->own) h->g_replays++;
        it += P;
        continue;
      }
      // capture (the host state advances as the pairs are recorded; it is
      // restored if the capture fails, since nothing has run)
      if (shifted && (rc = set_shift(0))) return rc;
      // a replay of the old graph may still be running
      if (b->gexec && hipStreamSynchronize(b->stream) != hipSuccess)
        return bfail(b, IRIS_EDEVICE, "hipStreamSynchronize failed");
      if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->gexec = nullptr;
      b->graph = nullptr;
      std::vector<std::pair<int, int>> saved;
      for (iris_handle* h : b->own) saved.push_back({h->nsamp, h->samp_n0});
      if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bfail(b, IRIS_EDEVICE, "hipStreamBeginCapture failed");
      int crc = 0;
      for (int k = 0; k < P && !crc; k += 2) crc = pair(it + k);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      hipGraphExec_t x = nullptr;
      hipError_t ei = hipSuccess;
      if (!crc && e == hipSuccess) ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (crc || e != hipSuccess || ei != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        pp = 0;
        for (size_t i = 0; i < saved.size(); ++i) {
          b->own[i]->nsamp = saved[i].first;
          b->own[i]->samp_n0 = saved[i].second;
        }
        if (crc) return crc;
        return bfail(b, IRIS_EDEVICE, "graph capture of batched steps %d..%d: %s", it, it + P - 1,
                     hipGetErrorString(e != hipSuccess ? e : ei));
      }
      b->graph = g;
      b->gexec = x;
      b->g_it0 = it;
      if (hipGraphLaunch(b->gexec, b->stream) != hipSuccess) return bfail(b, IRIS_EDEVICE, "hipGraphLaunch failed");
      it += P;
      continue;
    }
    if (shifted && (rc = set_shift(0))) return rc;
    if ((rc = pair(it))) return rc;
    it += 2;
  }
  if (shifted && (rc = set_shift(0))) return rc;
  for (iris_handle* h : b->own) {
    if (pp) {
      std::swap(h->fe, h->ke);
      std::swap(h->fh, h->kh);
      std::swap(h->d, h->kd);
      std::swap(h->b, h->kb);
    }
    h->it = it - 1;
  }
  for (iris_handle* h : b->own)
    if ((rc = run_steps<T>(h, it, it_last))) {
      b->err = h->err;
      return rc;
    }
  return 0;
}

// The batch's layout of a handle.  Alone, a small grid takes short z-chunks
// so that its launches have enough workgroups, and leaves its UPML slabs to
// single steps (a launch of the UPML variant over a few dozen workgroups
// took longer than the single steps).  In a batch the launches span every
// handle, so the z-chunks are those of a large grid (64-plane two-step
// chunks, 32-plane single-step ones; the chunk length changes no result) and
// the UPML slabs go into the two-step pass's UPML variant wherever it applies
// (bitwise the single steps without a Bloch phase; with one, equal to
// rounding, as the two-step pass is).  IRIS_BATCH_UPML=0 keeps what each
// handle alone does there, so that every result is bitwise that of the
// handle alone.
PassConfig batch_layout(const iris_handle* h) {
  PassConfig c = h->pass;
  c.zc = std::min(32, h->p.nz);
  c.zc2 = 64;
  const char* e = std::getenv("IRIS_BATCH_UPML");
  c.upml = (e && e[0] == '0') ? (h->n2p > 0 ? 1 : 0) : 1;
  return c;
}

template <typename T>
int batch_create_impl(iris_batch* b) {
  int rc;
  for (iris_handle* h : b->hs) {
    b->saved.push_back(h->pass);
    if ((rc = build_passes<T>(h, batch_layout(h)))) {
      b->err = h->err;
      return rc;
    }
    b->pass_gen.push_back(h->pass_gen);
  }
  // handles the two-step pass covers keep it (their own passes, as alone);
  // the batched single-step pass takes the others
  for (iris_handle* h : b->hs) (h->two_step ? b->own : b->one).push_back(h);
  const int n = (int)b->one.size();
  // z-chunks as long as the whole batch still gives every CU many workgroups
  int64_t tiles = 0;
  int nzmax = 0;
  for (iris_handle* h : b->one) {
    tiles += (int64_t)h->grd_fused.x * h->grd_fused.y;
    nzmax = std::max(nzmax, h->p.nz);
  }
  int zc = 32;
  while (zc > 1 && tiles * ((nzmax + zc - 1) / zc) < 16 * 256) zc /= 2;
  const char* zenv = std::getenv("IRIS_BATCH_ZCHUNK");
  if (zenv && std::atoi(zenv) > 0) zc = std::atoi(zenv);
  b->zc = zc;
  std::vector<int> wg[2] = {{0}, {0}};
  std::vector<int64_t> start = {0};
  for (int i = 0; i < n; ++i) {
    iris_handle* h = b->one[i];
    std::vector<int> lean, gen;
    single_lists(h->p, std::min(zc, h->p.nz), lean, gen);
    int* dl = nullptr;
    int* dg = nullptr;
    if ((rc = balloc(b, (void**)&dl, sizeof(int) * std::max<size_t>(lean.size(), 1))) ||
        (rc = balloc(b, (void**)&dg, sizeof(int) * std::max<size_t>(gen.size(), 1))))
      return rc;
    if (!lean.empty()) BHIP_TRY(b, hipMemcpy(dl, lean.data(), sizeof(int) * lean.size(), hipMemcpyHostToDevice));
    if (!gen.empty()) BHIP_TRY(b, hipMemcpy(dg, gen.data(), sizeof(int) * gen.size(), hipMemcpyHostToDevice));
    b->lean.push_back(dl);
    b->gen.push_back(dg);
    b->nlean.push_back((int)lean.size());
    b->ngen.push_back((int)gen.size());
    const int t = (int)(h->grd_fused.x * h->grd_fused.y);
    wg[0].push_back(wg[0].back() + t * (int)lean.size());
    wg[1].push_back(wg[1].back() + t * (int)gen.size());
    for (int bx = 0; bx < h->p.nbox; ++bx)
      if (h->box[bx].g.ntot) start.push_back(start.back() + (h->box[bx].g.ntot + 255) / 256 * 256);
  }
  for (int k = 0; k < 2; ++k) {
    if ((rc = balloc(b, (void**)&b->wg[k], sizeof(int) * wg[k].size()))) return rc;
    BHIP_TRY(b, hipMemcpy(b->wg[k], wg[k].data(), sizeof(int) * wg[k].size(), hipMemcpyHostToDevice));
    b->nwg[k] = wg[k].back();
  }
  b->nent = (int)start.size() - 1;
  b->npts = start.back();
  if ((rc = balloc(b, (void**)&b->ent_start, sizeof(int64_t) * start.size())) ||
      (rc = balloc(b, &b->inst, sizeof(BatchInst<T>) * std::max(n, 1))) ||
      (rc = balloc(b, &b->ent, sizeof(BatchSamp<T>) * std::max(b->nent, 1))))
    return rc;
  BHIP_TRY(b, hipMemcpy(b->ent_start, start.data(), sizeof(int64_t) * start.size(), hipMemcpyHostToDevice));
  return batch_create2<T>(b);
}

template <typename T>
int batch_step_impl(iris_batch* b, int it_first, int it_last) {
  const int n = (int)b->one.size();
  const int nall = (int)b->hs.size();
  int rc;
  // the instances' own work and the batch's previous run first (the argument
  // tables are rewritten); the instances' launches then go to the batch stream
  BHIP_TRY(b, hipStreamSynchronize(b->stream));
  for (iris_handle* h : b->hs) BHIP_TRY(b, hipStreamSynchronize(h->stream));
  if ((rc = batch_tables<T>(b)) || (rc = batch_tables2<T>(b))) return rc;
  std::vector<hipStream_t> own(nall);
  for (int i = 0; i < nall; ++i) {
    own[i] = b->hs[i]->stream;
    b->hs[i]->stream = b->stream;
  }
  rc = 0;
  // the two-step handles: batched pairs of steps, on the batch's stream
  rc = batch_run2<T>(b, it_first, it_last);
  iris_handle* h0 = n ? b->one[0] : nullptr;
  int par = 0;
  for (int it = it_first; n && it <= it_last && !rc; ++it) {
    for (int k = 0; k < 2; ++k) {
      if (!b->nwg[k]) continue;
      if (k == 0)
        hipLaunchKernelGGL((k_yee_fused_batch<T, false>), dim3(b->nwg[0]), dim3(kFTX, kFTY + 1), 0, b->stream,
                           (const BatchInst<T>*)b->inst, (const int*)b->wg[0], n, par, it);
      else
        hipLaunchKernelGGL((k_yee_fused_batch<T, true>), dim3(b->nwg[1]), dim3(kFTX, kFTY + 1), 0, b->stream,
                           (const BatchInst<T>*)b->inst, (const int*)b->wg[1], n, par, it);
    }
    par ^= 1;
    if (hipGetLastError() != hipSuccess) {
      rc = bfail(b, IRIS_EDEVICE, "batched pass launch failed");
      break;
    }
    if (it >= h0->p.to_fourier && (it % h0->p.sample_fourier) == 0 && b->nent) {
      // the step just taken is in the buffers of parity `par`'s source side
      hipLaunchKernelGGL(k_dft_sample_batch<T>, dim3((unsigned)((b->npts + 255) / 256)), dim3(256), 0, b->stream,
                         (const BatchSamp<T>*)b->ent, (const int64_t*)b->ent_start, b->nent, par, h0->nsamp, 0);
      for (iris_handle* h : b->one) {
        if (h->nsamp == 0) h->samp_n0 = it + 1 - h->p.to_fourier;
        h->nsamp++;
        if (h->nsamp == dft_slots<T>() && (rc = dft_flush<T>(h))) {
          b->err = h->err;
          break;
        }
      }
    }
  }
  const int steps = par;  // parity of the number of steps taken
  for (int i = 0; i < n; ++i) {
    iris_handle* h = b->one[i];
    if (steps) {
      std::swap(h->fe, h->ge);
      std::swap(h->fh, h->gh);
      std::swap(h->d, h->dn);
      std::swap(h->b, h->bn);
    }
    if (!rc) h->it = it_last;
  }
  for (int i = 0; i < nall; ++i) b->hs[i]->stream = own[i];
  if (rc) return rc;
  // later work on each instance's own stream waits for the batch
  BHIP_TRY(b, hipEventRecord(b->done, b->stream));
  for (iris_handle* h : b->hs) BHIP_TRY(b, hipStreamWaitEvent(h->stream, b->done, 0));
  return 0;
}

void batch_release(iris_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  for (void* q : b->allocs) (void)hipFree(q);
  b->allocs.clear();
  if (b->done) (void)hipEventDestroy(b->done);
  b->done = nullptr;
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

// The handles' own layouts back, after the batch's work has finished (a
// handle whose layout was rebuilt since the batch took it keeps the new one).
int batch_restore(iris_batch* b) {
  int rc = 0;
  for (size_t i = 0; i < b->saved.size(); ++i) {
    iris_handle* h = b->hs[i];
    if (i < b->pass_gen.size() && h->pass_gen != b->pass_gen[i]) continue;
    const int r = h->prec == 8 ? build_passes<double>(h, b->saved[i]) : build_passes<float>(h, b->saved[i]);
    if (r && !rc) {
      rc = r;
      b->err = h->err;
    }
  }
  return rc;
}

void release(iris_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  graph_release(h);
  prof_drain(h);
  for (hipEvent_t e : h->pool) (void)hipEventDestroy(e);
  h->pool.clear();
  for (void* p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  for (void* p : h->ck) (void)hipFree(p);
  h->ck.clear();
  h->ck_ok.clear();
  if (h->tail_rec) (void)hipFree(h->tail_rec);
  h->tail_rec = nullptr;
  h->ck_final = nullptr;  // one of adj_allocs
  for (void* p : h->adj_allocs) (void)hipFree(p);
  h->adj_allocs.clear();
  if (h->diag_buf) (void)hipFree(h->diag_buf);
  h->diag_buf = nullptr;
  if (h->plane_buf) (void)hipFree(h->plane_buf);
  h->plane_buf = nullptr;
  h->plane_cap = 0;
  if (h->stream) (void)hipStreamDestroy(h->stream);
  h->stream = nullptr;
}

// field index of (ix,iy,iz) in device layout
inline int64_t dev_index(const iris_handle* h, int ix, int iy, int iz) {
  return (int64_t)(ix + h->xoff) + h->sx * iy + h->sxy * iz;
}

}  // namespace

extern "C" {

int iris_create(const iris_params* p, const double* eps_inv, const double* fpzmn, const double* fmzmn,
                const double* fpzmx, const double* fmzmx, const double* fw, iris_handle** out) {
  if (!out) return fail(nullptr, IRIS_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = check_params(p);
  if (rc) return rc;
  if (!eps_inv || !fpzmn || !fmzmn || !fpzmx || !fmzmx || !fw) return fail(nullptr, IRIS_EINVAL, "NULL array argument");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) return fail(nullptr, IRIS_EDEVICE, "no HIP device available");
  iris_handle* h = new iris_handle();
  h->p = *p;
  h->prec = p->precision;
  {
    const char* env = std::getenv("IRIS_FUSED");
    h->fused = !(env && env[0] == '0');
    const char* aenv = std::getenv("IRIS_ADJ_FUSED");
    h->adj_fused = !(aenv && aenv[0] == '0');
    const char* tyenv = std::getenv("IRIS_ADJ_TY");
    if (tyenv && (std::atoi(tyenv) == 7 || std::atoi(tyenv) == 8 || std::atoi(tyenv) == 15))
      h->adj_ty = std::atoi(tyenv);
    // Graphs are off under a profiler's tool library (rocprofv3 sets
    // ROCP_TOOL_LIBRARIES): a kernel trace of many handles each replaying
    // its own graph ended in a host-side fault inside the HIP runtime with
    // no fault of a kernel; IRIS_GRAPH=1 keeps them on regardless.
    const char* genv = std::getenv("IRIS_GRAPH");
    const char* tool = std::getenv("ROCP_TOOL_LIBRARIES");
    h->graph_on = genv ? genv[0] != '0' : !(tool && tool[0]);
    const char* fp = std::getenv("IRIS_FOLD_PERSIST");
    if (fp && *fp) h->fold_persist = std::max(0, std::atoi(fp));
  }
  (void)hipGetDevice(&h->device);
  if (hipDeviceGetAttribute(&h->ncu, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || h->ncu < 1)
    h->ncu = 256;
  e = scratch_setup(h->device);
  if (e != hipSuccess) {
    delete h;
    return fail(nullptr, IRIS_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(nullptr, IRIS_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  rc = (h->prec == 8) ? create_impl<double>(h, eps_inv, fpzmn, fmzmn, fpzmx, fmzmx, fw)
                      : create_impl<float>(h, eps_inv, fpzmn, fmzmn, fpzmx, fmzmx, fw);
  if (rc) {
    g_create_error = h->err;
    release(h);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

int iris_step(iris_handle* h, int32_t it_first, int32_t it_last) {
  if (!h) return fail(nullptr, IRIS_EINVAL, "handle is NULL");
  if (it_first != h->it + 1) return fail(h, IRIS_ESTATE, "it_first=%d but the next step is %d", it_first, h->it + 1);
  if (it_last < it_first) return 0;
  if (it_last > h->p.nt) return fail(h, IRIS_ESTATE, "it_last=%d is past nt=%d", it_last, h->p.nt);
  (void)hipSetDevice(h->device);
  if (h->ck_seg > 0) return step_with_checkpoints(h, it_first, it_last);
  return h->prec == 8 ? run_steps<double>(h, it_first, it_last) : run_steps<float>(h, it_first, it_last);
}

int iris_checkpoint(iris_handle* h, int32_t segment, int64_t record_bytes) {
  if (!h) return fail(nullptr, IRIS_EINVAL, "handle is NULL");
  if (h->it != 0) return fail(h, IRIS_ESTATE, "iris_checkpoint: call it before the first step (now at step %d)", h->it);
  (void)hipSetDevice(h->device);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (void* q : h->ck) (void)hipFree(q);
  h->ck.clear();
  h->ck_ok.clear();
  if (h->tail_rec) (void)hipFree(h->tail_rec);
  h->tail_rec = nullptr;
  h->tail_ok.clear();
  h->ck_seg = 0;
  h->tail_t0 = 0;
  h->tail_n = 0;
  if (segment == 0) return 0;
  const iris_params& p = h->p;
  const int nt = p.nt;
  h->rec_plane = h->sxy * (p.hf - p.hi + 2);
  const size_t rstep = rec_step_bytes(h), sbytes = state_bytes(h);
  if (segment < 0) {
    // automatic: the snapshots ((nt / segment) states) and one segment's
    // records (segment records of the structure planes) take the least
    // memory together at segment = sqrt(nt * state / record); never below
    // ceil(sqrt(nt)).  On a 256 x 256 x 1024 grid with N = 16000 that is
    // ~1000 steps: 16 snapshots instead of 126 (which would not fit).
    const double s_mem = std::sqrt((double)nt * (double)sbytes / (double)rstep);
    const double s_min = std::ceil(std::sqrt((double)nt));
    segment = (int32_t)std::min<double>(nt, std::max(s_min, std::floor(s_mem)));
    segment = std::max<int32_t>(segment, 1);
  }
  int64_t nrec = 0;
  if (record_bytes > 0) {
    nrec = std::min<int64_t>(nt, record_bytes / (int64_t)rstep);
  } else if (record_bytes < 0) {
    // as many recorded steps as fit in 90% of the free device memory beside
    // the head snapshots and the adjoint's own buffers
    size_t freeb = 0, total = 0;
    HIP_TRY(h, hipMemGetInfo(&freeb, &total));
    const size_t fb = field_bytes(h), ab = aux_bytes(h);
    int64_t ntot = 0;
    for (int b = 0; b < p.nbox; ++b) ntot = std::max<int64_t>(ntot, h->box[b].g.ntot);
    const double reserve = 7.0 * fb + 4.0 * ab + (double)segment * rstep + sbytes + 16.0 * ntot * p.nof +
                           16.0 * p.nx * p.ny * (p.hf - p.hi + 1) + (double)(1u << 30);
    const double avail = 0.9 * (double)freeb - reserve;
    nrec = std::min<int64_t>(nt, avail > 0 ? (int64_t)(avail / rstep) : 0);
    while (nrec > 0 && (double)nrec * rstep + (double)((nt - nrec + segment - 1) / segment) * sbytes > avail)
      nrec -= std::max<int64_t>(1, nrec / 64);
    nrec = std::max<int64_t>(nrec, 0);
  }
  const int t0 = nt - (int)nrec;
  const int nhead = (t0 + segment - 1) / segment;
  if (nrec > 0 && hipMalloc(&h->tail_rec, rstep * (size_t)nrec) != hipSuccess) {
    (void)hipGetLastError();
    h->tail_rec = nullptr;
    return fail(h, IRIS_ENOMEM, "iris_checkpoint: records of %lld steps (%zu bytes each) do not fit", (long long)nrec,
                rstep);
  }
  for (int k = 0; k < nhead; ++k) {
    void* q = nullptr;
    if (hipMalloc(&q, sbytes) != hipSuccess) {
      (void)hipGetLastError();
      for (void* r : h->ck) (void)hipFree(r);
      h->ck.clear();
      if (h->tail_rec) (void)hipFree(h->tail_rec);
      h->tail_rec = nullptr;
      return fail(h, IRIS_ENOMEM, "iris_checkpoint: %d snapshots of %zu bytes do not fit", nhead, sbytes);
    }
    h->ck.push_back(q);
  }
  h->ck_ok.assign(h->ck.size(), 0);
  h->tail_ok.assign((size_t)nrec, 0);
  h->tail_t0 = t0;
  h->tail_n = (int)nrec;
  h->ck_seg = segment;
  if (h->two_step) {
    const int rc = h->prec == 8 ? build_lists<double>(h, h->rl, 1, 0, true) : build_lists<float>(h, h->rl, 1, 0, true);
    if (rc) return rc;
  }
  return 0;
}

int iris_checkpoint_info(const iris_handle* h, int64_t info[6]) {
  if (!h || !info) return IRIS_EINVAL;
  info[0] = h->ck_seg;
  info[1] = (int64_t)h->ck.size();
  info[2] = h->tail_n;
  info[3] = h->tail_t0;
  info[4] = (int64_t)(h->csize * 3 * (size_t)h->sxy * (h->p.hf - h->p.hi + 2));
  info[5] = (int64_t)state_bytes(h);
  return 0;
}

int iris_adjoint(iris_handle* h, int32_t nobox, const double* weights, double* grad) {
  if (!h || !weights || !grad) return fail(h, IRIS_EINVAL, "NULL argument");
  if (nobox < 1 || nobox > h->p.nbox) return fail(h, IRIS_EINVAL, "box %d does not exist", nobox);
  if (h->ck_seg <= 0) return fail(h, IRIS_ESTATE, "iris_adjoint: no snapshots (call iris_checkpoint before the first step)");
  if (h->it != h->p.nt) return fail(h, IRIS_ESTATE, "iris_adjoint: the forward run is at step %d, not nt=%d", h->it, h->p.nt);
  for (size_t k = 0; k < h->ck_ok.size(); ++k)
    if (!h->ck_ok[k]) return fail(h, IRIS_ESTATE, "iris_adjoint: snapshot %zu was not taken", k);
  for (size_t k = 0; k < h->tail_ok.size(); ++k)
    if (!h->tail_ok[k]) return fail(h, IRIS_ESTATE, "iris_adjoint: H of step %zu was not recorded", h->tail_t0 + k);
  (void)hipSetDevice(h->device);
  return h->prec == 8 ? adjoint_impl<double>(h, nobox, weights, grad) : adjoint_impl<float>(h, nobox, weights, grad);
}

int iris_sync(iris_handle* h) {
  if (!h) return fail(nullptr, IRIS_EINVAL, "handle is NULL");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

int iris_flux(iris_handle* h, int32_t nobox, double* jabs) {
  if (!h || !jabs) return fail(h, IRIS_EINVAL, "NULL argument");
  if (nobox < 1 || nobox > h->p.nbox) return fail(h, IRIS_EINVAL, "box %d does not exist", nobox);
  int rc = flush(h);
  if (rc) return rc;
  DeviceBox& B = h->box[nobox - 1];
  const int nof = h->p.nof;
  HIP_TRY(h, hipMemsetAsync(h->flux, 0, sizeof(double) * 6 * nof, h->stream));
  {
    ProfScope ps(h, IRIS_K_FLUX);
    hipLaunchKernelGGL(k_flux, dim3(6, nof), dim3(256), 0, h->stream, B.g, (const double2*)B.emg, nof,
                       B.active_mask, h->flux);
  }
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(jabs, h->flux, sizeof(double) * 6 * nof, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  // Jabs = (eps0*q/(2.0*dt))*Jabs, reference code/dataprocess.f90:277
  const double s = (h->p.eps0 * h->p.q) / (2.0 * h->p.dt);
  for (int i = 0; i < 6 * nof; ++i) jabs[i] = s * jabs[i];
  return 0;
}

int iris_dft_points(const iris_handle* h, int32_t nobox, int32_t pair, int64_t* npts) {
  if (!h || !npts) return fail(nullptr, IRIS_EINVAL, "NULL argument");
  if (nobox < 1 || nobox > h->p.nbox || pair < 1 || pair > 3) return fail(nullptr, IRIS_EINVAL, "bad box or pair");
  *npts = h->box[nobox - 1].npts_ref[pair - 1];
  return 0;
}

int iris_get_dft(iris_handle* h, int32_t nobox, int32_t pair, double* out) {
  if (!h || !out) return fail(h, IRIS_EINVAL, "NULL argument");
  if (nobox < 1 || nobox > h->p.nbox || pair < 1 || pair > 3) return fail(h, IRIS_EINVAL, "bad box or pair");
  int rc = flush(h);
  if (rc) return rc;
  DeviceBox& B = h->box[nobox - 1];
  const int j = pair - 1;
  const size_t n = (size_t)B.npts_ref[j] * h->p.nof;
  if (B.g.npts[j] == 0) {
    std::memset(out, 0, sizeof(double2) * n);
    return 0;
  }
  HIP_TRY(h, hipMemcpyAsync(out, B.emg + B.g.start[j] * h->p.nof, sizeof(double2) * n, hipMemcpyDeviceToHost,
                            h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

int iris_get_field(iris_handle* h, int32_t which, double* out) {
  if (!h || !out || which < 0 || which > 1) return fail(h, IRIS_EINVAL, "bad argument");
  const iris_params& p = h->p;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  std::vector<char> buf(h->csize * (size_t)h->ncell);
  const int64_t rsx = p.nx + 2, rsxy = rsx * (p.ny + 2);
  for (int c = 0; c < 3; ++c) {
    const char* src = (const char*)(which ? h->fh : h->fe) + h->csize * (size_t)h->cs * c;
    HIP_TRY(h, hipMemcpyAsync(buf.data(), src, buf.size(), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int z = 0; z <= p.nz + 1; ++z)
      for (int y = 0; y <= p.ny + 1; ++y)
        for (int x = 0; x <= p.nx + 1; ++x) {
          const int64_t di = dev_index(h, x, y, z);
          double* o = out + 2 * (c + 3 * (x + rsx * y + rsxy * z));
          if (h->prec == 8) {
            const double2 v = ((const double2*)buf.data())[di];
            o[0] = v.x; o[1] = v.y;
          } else {
            const float2 v = ((const float2*)buf.data())[di];
            o[0] = v.x; o[1] = v.y;
          }
        }
  }
  return 0;
}

int iris_set_field(iris_handle* h, int32_t which, const double* in) {
  if (!h || !in || which < 0 || which > 1) return fail(h, IRIS_EINVAL, "bad argument");
  const iris_params& p = h->p;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  std::vector<char> buf(h->csize * (size_t)h->ncell, 0);
  const int64_t rsx = p.nx + 2, rsxy = rsx * (p.ny + 2);
  for (int c = 0; c < 3; ++c) {
    for (int z = 0; z <= p.nz + 1; ++z)
      for (int y = 0; y <= p.ny + 1; ++y)
        for (int x = 0; x <= p.nx + 1; ++x) {
          const int64_t di = dev_index(h, x, y, z);
          const double* v = in + 2 * (c + 3 * (x + rsx * y + rsxy * z));
          if (h->prec == 8) ((double2*)buf.data())[di] = make_double2(v[0], v[1]);
          else ((float2*)buf.data())[di] = make_float2((float)v[0], (float)v[1]);
        }
    char* dst = (char*)(which ? h->fh : h->fe) + h->csize * (size_t)h->cs * c;
    HIP_TRY(h, hipMemcpyAsync(dst, buf.data(), buf.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  return 0;
}

int iris_get_plane(iris_handle* h, int32_t what, int32_t axis, int32_t index, const int32_t range[4],
                   double* out) {
  if (!h || !range || !out) return fail(h, IRIS_EINVAL, "bad argument");
  const iris_params& p = h->p;
  if (what < 0 || what > 8) return fail(h, IRIS_EINVAL, "iris_get_plane: what must be 0..8");
  if (axis < 1 || axis > 3) return fail(h, IRIS_EINVAL, "iris_get_plane: axis must be 1, 2 or 3");
  const int n[3] = {p.nx, p.ny, p.nz};
  const int fa = axis == 1 ? 1 : 0, fb = axis == 3 ? 1 : 2;  // free axes
  if (index < 1 || index > n[axis - 1] || range[0] < 1 || range[0] > range[1] || range[1] > n[fa] ||
      range[2] < 1 || range[2] > range[3] || range[3] > n[fb])
    return fail(h, IRIS_EINVAL, "iris_get_plane: index %d or range [%d,%d]x[%d,%d] outside 1..n of axis %d", index,
                range[0], range[1], range[2], range[3], axis);
  const int base = what / 3, comp = what % 3;
  const void* f;
  if (base == 0) f = h->fe;
  else if (base == 2) f = h->fh;
  else if (h->fused) f = h->gh;  // the fused pass read H(it-1) from here and wrote H(it) into fh
  else return fail(h, IRIS_EUNSUP, "iris_get_plane: H of the previous step is kept only by the fused pass");
  const int na = range[1] - range[0] + 1, nb = range[3] - range[2] + 1;
  const int64_t np = (int64_t)na * nb;
  if (np > h->plane_cap) {
    if (h->plane_buf) {
      HIP_TRY(h, hipStreamSynchronize(h->stream));
      (void)hipFree(h->plane_buf);
      h->plane_buf = nullptr;
      h->plane_cap = 0;
    }
    if (hipMalloc(&h->plane_buf, h->csize * (size_t)np) != hipSuccess) {
      (void)hipGetLastError();
      return fail(h, IRIS_ENOMEM, "iris_get_plane: %lld complex values", (long long)np);
    }
    h->plane_cap = np;
  }
  const char* src = (const char*)f + h->csize * (size_t)h->cs * comp;
  const dim3 grid((unsigned)((np + 255) / 256));
  if (h->prec == 8)
    hipLaunchKernelGGL(k_plane<double>, grid, dim3(256), 0, h->stream, (const double2*)src, axis, index, range[0], na,
                       range[2], nb, h->sx, h->sxy, h->xoff, (double2*)h->plane_buf);
  else
    hipLaunchKernelGGL(k_plane<float>, grid, dim3(256), 0, h->stream, (const float2*)src, axis, index, range[0], na,
                       range[2], nb, h->sx, h->sxy, h->xoff, (float2*)h->plane_buf);
  HIP_TRY(h, hipGetLastError());
  if (h->prec == 8) {
    HIP_TRY(h, hipMemcpyAsync(out, h->plane_buf, sizeof(double2) * (size_t)np, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  } else {
    std::vector<float2> buf((size_t)np);
    HIP_TRY(h, hipMemcpyAsync(buf.data(), h->plane_buf, sizeof(float2) * (size_t)np, hipMemcpyDeviceToHost,
                              h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int64_t i = 0; i < np; ++i) {
      out[2 * i] = buf[(size_t)i].x;
      out[2 * i + 1] = buf[(size_t)i].y;
    }
  }
  return 0;
}

int iris_get_plane_noalloc(iris_handle* h, int32_t what, int32_t axis, int32_t index, const int32_t range[4],
                           double* out) {
  if (!h || !range || !out) return fail(h, IRIS_EINVAL, "bad argument");
  const iris_params& p = h->p;
  if (what < 0 || what > 8) return fail(h, IRIS_EINVAL, "iris_get_plane_noalloc: what must be 0..8");
  if (axis < 1 || axis > 3) return fail(h, IRIS_EINVAL, "iris_get_plane_noalloc: axis must be 1, 2 or 3");
  const int n[3] = {p.nx, p.ny, p.nz};
  const int fa = axis == 1 ? 1 : 0, fb = axis == 3 ? 1 : 2;  // free axes
  if (index < 1 || index > n[axis - 1] || range[0] < 1 || range[0] > range[1] || range[1] > n[fa] ||
      range[2] < 1 || range[2] > range[3] || range[3] > n[fb])
    return fail(h, IRIS_EINVAL, "iris_get_plane_noalloc: index %d or range [%d,%d]x[%d,%d] outside 1..n of axis %d",
                index, range[0], range[1], range[2], range[3], axis);
  const int na = range[1] - range[0] + 1, nb = range[3] - range[2] + 1;
  const int64_t np = (int64_t)na * nb;
  if (np > kPlaneScratch)
    return fail(h, IRIS_EINVAL, "iris_get_plane_noalloc: %lld values exceed the %lld of its fixed staging array "
                "(use iris_get_plane)", (long long)np, (long long)kPlaneScratch);
  const int base = what / 3, comp = what % 3;
  const void* f;
  if (base == 0) f = h->fe;
  else if (base == 2) f = h->fh;
  else if (h->fused) f = h->gh;  // the fused pass read H(it-1) from here and wrote H(it) into fh
  else return fail(h, IRIS_EUNSUP, "iris_get_plane_noalloc: H of the previous step is kept only by the fused pass");
  void* stage = g_scratch_ptr[h->device];  // resolved at iris_create on this device
  if (!stage) return fail(h, IRIS_ESTATE, "iris_get_plane_noalloc: staging array of device %d unknown", h->device);
  const char* src = (const char*)f + h->csize * (size_t)h->cs * comp;
  const dim3 grid((unsigned)((np + 255) / 256));
  if (h->prec == 8)
    hipLaunchKernelGGL(k_plane<double>, grid, dim3(256), 0, h->stream, (const double2*)src, axis, index, range[0], na,
                       range[2], nb, h->sx, h->sxy, h->xoff, (double2*)stage);
  else
    hipLaunchKernelGGL(k_plane<float>, grid, dim3(256), 0, h->stream, (const float2*)src, axis, index, range[0], na,
                       range[2], nb, h->sx, h->sxy, h->xoff, (float2*)stage);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(out, stage, h->csize * (size_t)np, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (h->prec == 4) {
    // widen the float2 values in place, from the last one down (value i
    // moves from bytes 8i to 16i, never over a value not yet moved)
    unsigned char* bytes = (unsigned char*)out;
    for (int64_t i = np - 1; i >= 0; --i) {
      float v[2];
      std::memcpy(v, bytes + 8 * i, sizeof v);
      const double w[2] = {v[0], v[1]};
      std::memcpy(bytes + 16 * i, w, sizeof w);
    }
  }
  return 0;
}

int iris_diagnostics(iris_handle* h, int32_t zi, int32_t zf, double out[4]) {
  if (!h || !out) return fail(h, IRIS_EINVAL, "bad argument");
  const iris_params& p = h->p;
  if (zi < p.pmlzi + 1 || zf > p.pmlzf - 1 || zi > zf)
    return fail(h, IRIS_EINVAL, "iris_diagnostics: z range [%d,%d] must lie in [pmlzi+1, pmlzf-1] = [%d,%d]", zi,
                zf, p.pmlzi + 1, p.pmlzf - 1);
  if (!h->fused) return fail(h, IRIS_EUNSUP, "iris_diagnostics: E and H of the previous step are kept only by the fused pass");
  if (!h->diag_buf) {
    if (hipMalloc((void**)&h->diag_buf, sizeof(double) * (4 * kDiagBlocks + 4)) != hipSuccess) {
      (void)hipGetLastError();
      h->diag_buf = nullptr;
      return fail(h, IRIS_ENOMEM, "iris_diagnostics: partials");
    }
  }
  double* res = h->diag_buf + 4 * kDiagBlocks;
  if (h->prec == 8) {
    const StepArgs<double> a = step_args<double>(h);
    hipLaunchKernelGGL(k_diag<double>, dim3(kDiagBlocks), dim3(256), 0, h->stream, a, (const double2*)h->ge,
                       (const double2*)h->gh, zi, zf, h->dtcq2, h->diag_buf);
  } else {
    const StepArgs<float> a = step_args<float>(h);
    hipLaunchKernelGGL(k_diag<float>, dim3(kDiagBlocks), dim3(256), 0, h->stream, a, (const float2*)h->ge,
                       (const float2*)h->gh, zi, zf, h->dtcq2, h->diag_buf);
  }
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, h->stream, (const double*)h->diag_buf, kDiagBlocks, res);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(out, res, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}

int iris_pass_info(const iris_handle* h, int64_t info[6]) {
  if (!h || !info) return IRIS_EINVAL;
  info[0] = h->zchunk;
  info[1] = h->two_step ? h->planes2 : 0;
  info[2] = h->two_step ? h->n2 : 0;
  info[3] = h->two_step ? 1 : 0;
  info[4] = h->two_step ? h->planes2p : 0;
  info[5] = h->two_step ? h->upml2p : 0;
  return 0;
}

int iris_pair_layout(int32_t nx, int32_t ny, int32_t only_ox, int32_t out[11]) {
  if (!out) return fail(nullptr, IRIS_EINVAL, "iris_pair_layout: out is NULL");
  const P2Layout L = p2_layout(nx, ny, only_ox);
  if (!L.ok) return fail(nullptr, IRIS_EINVAL, "iris_pair_layout: no layout of two cells a lane fits %d x %d", nx, ny);
  const int32_t v[11] = {L.ox, L.sl, L.ns, L.tx, L.ty, L.oxn, L.sln, L.nsn, L.txn, L.tyn, L.x0n};
  std::copy(v, v + 11, out);
  return 0;
}

int iris_graph_info(const iris_handle* h, int64_t info[3]) {
  if (!h || !info) return fail(nullptr, IRIS_EINVAL, "handle or info is NULL");
  info[0] = h->gexec ? 1 : 0;
  info[1] = h->g_replays;
  info[2] = (int64_t)(h->prec == 8 ? kDftSlots64 : kDftSlots32) * h->p.sample_fourier;
  return 0;
}

int iris_autotune(iris_handle* h, int32_t steps, int64_t info[9]) {
  if (!h || !info) return fail(h, IRIS_EINVAL, "handle or info is NULL");
  if (h->it != 0) return fail(h, IRIS_ESTATE, "iris_autotune: call it before the first step (the handle is at step %d)", h->it);
  if (h->ck_seg > 0) return fail(h, IRIS_ESTATE, "iris_autotune: call it before iris_checkpoint");
  if (!h->fused) return fail(h, IRIS_EUNSUP, "iris_autotune: the in-place passes have no layout to tune");
  if (steps < 2 || steps > 1000) return fail(h, IRIS_EINVAL, "iris_autotune: steps must be in 2..1000 (got %d)", steps);
  (void)hipSetDevice(h->device);
  return h->prec == 8 ? autotune_impl<double>(h, steps, info) : autotune_impl<float>(h, steps, info);
}

int iris_current_step(const iris_handle* h, int32_t* it) {
  if (!h || !it) return IRIS_EINVAL;
  *it = h->it;
  return 0;
}

int iris_destroy(iris_handle* h) {
  if (!h) return 0;
  release(h);
  delete h;
  return 0;
}

int iris_profile(iris_handle* h, int32_t enable) {
  if (!h) return fail(nullptr, IRIS_EINVAL, "handle is NULL");
  if (!enable) prof_drain(h);
  h->prof = enable != 0;
  return 0;
}

int iris_profile_read(iris_handle* h, double* ms, int64_t* launches, int32_t reset) {
  if (!h || !ms || !launches) return fail(h, IRIS_EINVAL, "NULL argument");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  prof_drain(h);
  for (int k = 0; k < IRIS_NKERNEL; ++k) {
    ms[k] = h->prof_ms[k];
    launches[k] = h->prof_n[k];
    if (reset) { h->prof_ms[k] = 0; h->prof_n[k] = 0; }
  }
  return 0;
}

const char* iris_last_error(const iris_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int iris_batch_create(iris_handle* const* handles, int32_t n, iris_batch** out) {
  if (!out) return bfail(nullptr, IRIS_EINVAL, "out is NULL");
  *out = nullptr;
  if (!handles || n < 1) return bfail(nullptr, IRIS_EINVAL, "iris_batch_create: no handles");
  for (int i = 0; i < n; ++i) {
    const iris_handle* h = handles[i];
    if (!h) return bfail(nullptr, IRIS_EINVAL, "iris_batch_create: handle %d is NULL", i);
    if (!h->fused) return bfail(nullptr, IRIS_EUNSUP, "iris_batch_create: handle %d runs the in-place passes", i);
    if (h->prec != handles[0]->prec || h->device != handles[0]->device)
      return bfail(nullptr, IRIS_EINVAL, "iris_batch_create: handle %d differs in precision or device", i);
    if (h->p.to_fourier != handles[0]->p.to_fourier || h->p.sample_fourier != handles[0]->p.sample_fourier)
      return bfail(nullptr, IRIS_EINVAL, "iris_batch_create: handle %d differs in the DFT cadence", i);
    if (h->ck_seg > 0) return bfail(nullptr, IRIS_EUNSUP, "iris_batch_create: handle %d keeps gradient snapshots", i);
    for (int j = 0; j < i; ++j)
      if (handles[j] == h) return bfail(nullptr, IRIS_EINVAL, "iris_batch_create: handle %d appears twice", i);
  }
  iris_batch* b = new iris_batch();
  b->hs.assign(handles, handles + n);
  b->prec = handles[0]->prec;
  (void)hipSetDevice(handles[0]->device);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    batch_release(b);
    delete b;
    return bfail(nullptr, IRIS_EDEVICE, "iris_batch_create: stream or event");
  }
  const int rc = b->prec == 8 ? batch_create_impl<double>(b) : batch_create_impl<float>(b);
  if (rc) {
    g_create_error = b->err;
    batch_release(b);
    batch_restore(b);
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int iris_batch_step(iris_batch* b, int32_t it_first, int32_t it_last) {
  if (!b) return bfail(nullptr, IRIS_EINVAL, "batch is NULL");
  for (size_t i = 0; i < b->hs.size(); ++i) {
    // first: a rebuilt layout (iris_autotune) leaves the batch's own/one
    // split and tables stale, so nothing below may rely on them
    if (b->hs[i]->pass_gen != b->pass_gen[i])
      return bfail(b, IRIS_ESTATE, "iris_batch_step: the pass layout of handle %zu changed after iris_batch_create "
                   "(iris_autotune); create the batch again", i);
  }
  for (size_t i = 0; i < b->hs.size(); ++i) {
    const iris_handle* h = b->hs[i];
    // the rings of each class (batched two-step handles, the others) fill together
    const bool own = std::find(b->own.begin(), b->own.end(), h) != b->own.end();
    const iris_handle* ring0 = own ? b->own[0] : b->one[0];
    if (h->it != it_first - 1)
      return bfail(b, IRIS_ESTATE, "iris_batch_step: handle %zu is at step %d, not %d", i, h->it, it_first - 1);
    if (it_last > h->p.nt) return bfail(b, IRIS_ESTATE, "iris_batch_step: it_last=%d is past nt=%d of handle %zu",
                                        it_last, h->p.nt, i);
    if (h->nsamp != ring0->nsamp) return bfail(b, IRIS_ESTATE, "iris_batch_step: DFT rings out of step");
    if (h->ck_seg > 0)
      return bfail(b, IRIS_EUNSUP, "iris_batch_step: handle %zu keeps gradient snapshots (step it alone)", i);
  }
  if (it_last < it_first) return 0;
  (void)hipSetDevice(b->hs[0]->device);
  return b->prec == 8 ? batch_step_impl<double>(b, it_first, it_last) : batch_step_impl<float>(b, it_first, it_last);
}

int iris_batch_info(const iris_batch* b, int64_t info[5]) {
  if (!b || !info) return IRIS_EINVAL;
  info[0] = (int64_t)b->hs.size();
  info[1] = b->zc;
  info[2] = b->nwg[0];
  info[3] = b->nwg[1];
  info[4] = (int64_t)b->own.size();
  return 0;
}

int iris_batch_destroy(iris_batch* b) {
  if (!b) return 0;
  batch_release(b);
  const int rc = batch_restore(b);
  delete b;
  return rc;
}

const char* iris_batch_last_error(const iris_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int iris_incident_table(const iris_params* p, double* out) {
  if (!p || !out) return fail(nullptr, IRIS_EINVAL, "iris_incident_table: NULL argument");
  if (!incident_matched(*p))
    return fail(nullptr, IRIS_EINVAL, "iris_incident_table: normal incidence takes the reference's analytic pulse");
  if (const char* why = incident_unsupported(*p)) return fail(nullptr, IRIS_EUNSUP, "iris_incident_table: %s", why);
  if (p->nt < 0 || p->nx < 1) return fail(nullptr, IRIS_EINVAL, "iris_incident_table: bad nt or nx");
  const std::vector<zcplx> tab = incident_table(*p);
  std::memcpy(out, tab.data(), sizeof(zcplx) * tab.size());
  return 0;
}

const char* iris_version(void) { return "iris-mi355x 0.2 (gfx950)"; }

}  // extern "C"
