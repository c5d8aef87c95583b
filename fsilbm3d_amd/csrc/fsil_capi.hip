// C-ABI of the FSILBM3D MI355X hot-path library: block lifetime, state
// transfer, the fused fluid step and the immersed-boundary coupling.
// Declarations and the reference routine each entry point replaces are in
// include/fsilbm3d_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fsilbm3d_amd.h"
#include "fsil_common.h"
#include "fsil_ibm_kernels.h"
#include "fsil_lbm_kernels.h"
#include "fsil_lbm32_kernels.h"
#include "fsil_fem_kernels.h"
#include "fsil_tree_kernels.h"
#include "fsil_mlp_kernels.h"

using namespace fsil;

namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return 1;
}

#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return fail("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                               \
  } while (0)

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // grows (never shrinks) the allocation; contents are not preserved
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = std::max<size_t>(count, 1);
    return e;
  }
};

// Host wait for an event of the step loop (the coupling call's report, the
// beam sub-steps): the event is polled for up to 5 ms before a blocking wait,
// whose wake-up latency would otherwise add to every step.  FSIL_SPIN=0 waits
// blocking from the start.
hipError_t step_wait(hipEvent_t ev) {
  static const bool spin = [] {
    const char *e = getenv("FSIL_SPIN");
    return !(e && !strcmp(e, "0"));
  }();
  if (spin) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      const hipError_t e = hipEventQuery(ev);
      if (e != hipErrorNotReady) return e;
      if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(5)) break;
    }
  }
  return hipEventSynchronize(ev);
}

// Pinned host copy of a coupling call's control words, written by k_ibm_finish.
struct PinnedReport {
  IbmReport *h = nullptr, *d = nullptr;  // host pointer and its device address
  hipEvent_t ev = nullptr;               // recorded after k_ibm_finish
  bool pending = false;                  // a report the host has not checked yet
  PinnedReport() = default;
  PinnedReport(const PinnedReport &) = delete;
  PinnedReport &operator=(const PinnedReport &) = delete;
  ~PinnedReport() {
    if (ev) (void)hipEventSynchronize(ev);
    if (ev) (void)hipEventDestroy(ev);
    if (h) (void)hipHostFree(h);
  }
  hipError_t ensure() {
    if (h) return hipSuccess;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) { ev = nullptr; return e; }
    e = hipHostMalloc((void **)&h, sizeof(IbmReport), hipHostMallocDefault);
    if (e != hipSuccess) { h = nullptr; return e; }
    memset(h, 0, sizeof(IbmReport));
    e = hipHostGetDevicePointer((void **)&d, h, 0);
    if (e != hipSuccess) { (void)hipHostFree(h); h = nullptr; d = nullptr; }
    return e;
  }
};

struct Ibm {
  int nbodies = 0;
  std::vector<int> nmark, nnode, nelem;      // per body
  std::vector<int> moff, noff, eoff;         // prefix offsets (markers, nodes, elements)
  int M = 0, nnodes = 0, nelems = 0;
  std::vector<int> marker_elem;              // global element id of each marker
  int ntol = 0;
  double dtol = 0.0, Uref = 1.0, denIn = 1.0;
  // device: markers and forces
  DevBuf<double> Exyz, Evel, Ea, Eforce, fT, tolpart, elem_xc, lod;
  DevBuf<int> dmarker_elem, dof_ptr, dof_list, dof_elem, dof_kind, first_marker;
  // device: stencils and band (built each call, see fsil_ibm_kernels.h)
  int cap = 0;                               // 64*M, stride of the band arrays
  DevBuf<float> dEw;                         // 12*M
  DevBuf<float4> went;                       // per band entry (CSR order): rx, ry, rz
  DevBuf<unsigned> keys;                      // cell of each stencil entry
  DevBuf<int> tmp;                            // entries in arrival order per band cell
  DevBuf<int> blkbox;                         // per claim workgroup: its part of the band's box
  DevBuf<int> flags, scan, dentry, dbidx, dptr, ctl, cellmap;
  DevBuf<long long> dcell;
  DevBuf<double> band_u, band_F, dstate;
  std::unique_ptr<PinnedReport> rep = std::make_unique<PinnedReport>();
  int ndof = 0;
  bool active = false;  // band valid for the next collision
  int box[6] = {0, 0, 0, 0, 0, 0};  // band box of the last checked call (kBoxLo..kBoxHi)
  bool have_box = false;
  int warm_mlp = 0;     // network handle of the learned warm start (0: off)
  int record = 0;       // compute the network's features and targets every call
  int skip = 0;         // forcing rounds the learned guess stands for
  bool have_prev = false;  // a coupling call has run since the bodies were defined (Eprev valid)
  DevBuf<float> wsX, wsY, wsP;               // features (12 per marker), targets, guesses (3 per marker)
  DevBuf<double> Eprev, Enrm, Esnap;         // previous force, normals, force after `skip` rounds
  int last_warm = 0, last_fallback = 0, last_rounds = 0;
  int last_iters = 0;
  int last_nband = 0;  // band cells of the last checked call (sizes the band-cell grids)
  int pass_cap = 0;    // entries per LDS pass of the band-cell sums (at most kPassEntries)
  double last_dmax = 0.0;
};

// Device FEM state of the bodies carried by a block.
struct FemBodyHost {
  int iBodyModel = 1;
  double Freq = 0, XYZo[3] = {0, 0, 0}, XYZAmpl[3] = {0, 0, 0}, XYZPhi[3] = {0, 0, 0}, AoAo[3] = {0, 0, 0},
         AoAAmpl[3] = {0, 0, 0}, AoAPhi[3] = {0, 0, 0}, initXYZVel[3] = {0, 0, 0};
  double TTT0[9];
};

// The beam sub-steps run on a stream of their own: the fluid update of a step
// does not read the beam state, so the sub-steps overlap the collide kernel.
// ev_in orders them after the coupling call; ev_done orders the next marker
// update (and the next coupling call) after them.  Descriptors and results
// travel through pinned host memory.
struct FemStream {
  hipStream_t s = nullptr;
  hipEvent_t ev_in = nullptr, ev_done = nullptr;
  FemBodyDesc *h_stage = nullptr;  // pinned: numsubstep x nbodies descriptors, read and
                                   // answered by the kernel in place
  FemBodyDesc *d_stage = nullptr;  // device address of h_stage
  size_t cap = 0;
  int n_staged = 0;                // descriptors of the last call (results after ev_done)
  bool pending = false;            // a call whose results have not been checked
  ~FemStream() {
    if (s) (void)hipStreamSynchronize(s);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (s) (void)hipStreamDestroy(s);
    if (h_stage) (void)hipHostFree(h_stage);
  }
};

struct Fem {
  int nbodies = 0, nEL = 0, nND = 0, gEQ = 0;
  std::unique_ptr<FemStream> fs;
  std::vector<FemBodyHost> bh;
  std::vector<FemBodyDesc> desc;
  double dampK = 0, dampM = 0, GeoGamma = 1, NewmarkGamma = 0.5, NewmarkBeta = 0.25, dtolFEM = 1e-10;
  int ntolFEM = 20;
  DevBuf<BeamElem> el;
  DevBuf<double> pos, dsp, vel, acc, dspO, velO, accO, lodInte, lodEffe, lodExte, vBC, x, r, p, Ap, z, xF, lbuf, blk,
      beta;
  DevBuf<int> fixed, dof_ptr, dof_e, dof_i, node_ptr, node_e, node_h, body_of_elem, elem_first_marker, body_goff;
  DevBuf<double> info;  // fsil_solid_info: per-element energies and the 78 row values
  bool markers_on_device = false;
  size_t lds_limit = 0;  // the device's LDS bytes per workgroup
  size_t lds_set[2] = {64 * 1024, 64 * 1024};  // dynamic LDS bytes k_fem_substep<false>, <true> may take
  // every element's rotated mass blocks repeat (the LDS element form holds
  // them); FSIL_FEM_HOT=0 keeps the element fields in global memory
  bool hot_ok = true;
};

// A stream shared by the blocks of one tree (father and sons run in order).
struct StreamHolder {
  hipStream_t s = nullptr;
  ~StreamHolder() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

// Father/son pair of the block tree (CommPair, LBMBlockComm.f90:11-18); the
// index arrays are kept 1-based as in the reference.
struct Pair {
  int son = 0;  // son block handle
  int sds[6], s[6], f[6], si[6], fi[6];
  int dimS[3], dimF[3];
  int scheme = 1;
  DevBuf<double> planes;  // per active face: level-1 and level-2 coarse planes
  long long off[6] = {0, 0, 0, 0, 0, 0};
  int extracted = 0;      // 2 once both levels hold planes of the same father step
};

struct Block {
  int id = 0;
  int xd = 0, yd = 0, zd = 0;
  long long N = 0;
  double dh = 1, xmin = 0, ymin = 0, zmin = 0, xmax = 0, ymax = 0, zmax = 0;
  int bc[6] = {0, 0, 0, 0, 0, 0};
  int model = 1;
  double tau = 1, Omega = 1, Omega2 = 1;
  int precision = 0;  // 0 FP64 plain, 1 FP32 deviation storage
  int nt = 0;         // population streams nontemporal (see stream_policy)
  int zrow = 0;       // FP32 row exchange of the z-moving values (zrow_exchange), fixed at creation
  bool nt_forced = false;  // FSIL_STREAM_HINT set at creation
  DevBuf<double> fA, fB;
  DevBuf<float> gA, gB;
  DevBuf<double> Mc, Mf;
  DevBuf<float> Mc32, Mf32;  // FP32-deviation MRT: the matrices rounded to FP32
  DevBuf<double> tau_all, uf;  // LES models: per-cell tau; the velocity field of models 14, 15
  DevBuf<float> outbuf;        // DatFlow fields of the last flow write (device staging)
  double nu = std::nan("");    // flow%nu (models 14, 15)
  // flow conditions (FlowCondition.f90:11-26)
  double denIn = 1, uvwIn[3] = {0, 0, 0}, shear[3] = {0, 0, 0};
  int vkind = 0;
  double VFin[3] = {0, 0, 0}, VFamp = 0, VFfreq = 0, VFphi = 0;
  double VF[3] = {0, 0, 0};
  double blktime = 0;
  // auxiliary planes
  DevBuf<double> nb_den[6], nb_u[6], hw[6];
  bool hw_valid[6] = {false, false, false, false, false, false};
  Ibm ibm;
  Fem fem;
  // block tree: the father's handle (0 for a root) and the son pairs
  int father = 0;
  std::vector<std::unique_ptr<Pair>> sons;
  std::shared_ptr<StreamHolder> sh;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // split steps (fsil_fsi_step): the coupling runs on the beam stream while
  // the bulk of the collision runs on the block's stream
  hipEvent_t ev_fluid = nullptr;  // the block's stream before the bulk collision
  bool split = false;             // the coupling's launches go to the beam stream
  bool split_ok = true;           // FSIL_SPLIT=0 turns split steps off
  int split_margin = 3;           // cells the predicted band box adds to the last one (FSIL_SPLIT_MARGIN)
  size_t bulk_lds = 0;            // dynamic LDS of an FP32 split step's bulk launch (its CU residency)
};

std::map<int, std::unique_ptr<Block>> g_blocks;
int g_next_handle = 1;
int g_device = -1;

// Device memory of fsil_solid_info_static, for callers that cannot allocate:
// the per-element energies (3 per element) and the 78 row values of one body
// of at most kStaticInfoElems elements.  Every device the library runs on
// holds one copy until the process ends; g_scratch keeps its device address
// per device (taken in fsil_init).  Calls on one device share the array, so
// calls that overlap on one device race.
constexpr int kStaticInfoElems = 256;
static __device__ __align__(256) double scratch[3 * kStaticInfoElems + 78];
std::map<int, double *> g_scratch;

// stream of the coupling's launches: the beam stream during a split step
hipStream_t coupling_stream(const Block &b) { return b.split ? b.fem.fs->s : b.stream; }

Block *get_block(int h) {
  auto it = g_blocks.find(h);
  if (it == g_blocks.end()) {
    fail("invalid block handle %d", h);
    return nullptr;
  }
  return it->second.get();
}

// the warm-start network (fsil_mlp_kernels.h), one per handle
struct Mlp {
  DevBuf<float> th, g, m, v;                         // parameters, gradient, Adam moments
  DevBuf<float> X, Y, bx, by, H1, H2, H3, Yp, dY, dZ3, dZ2, dZ1;
  DevBuf<double> part;
  DevBuf<int> idx;
  int adam_t = 0;
  hipStream_t s = nullptr;
  ~Mlp() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

std::map<int, std::unique_ptr<Mlp>> g_mlps;
int g_next_mlp = 1;

Mlp *get_mlp(int h) {
  auto it = g_mlps.find(h);
  if (it == g_mlps.end()) {
    fail("invalid network handle %d", h);
    return nullptr;
  }
  return it->second.get();
}

bool is_les(int model) { return model == 11 || model == 14 || model == 15; }

// ConstParams.f90:40-45 (Csmag 0.17, CWALE 0.5), in the reference's order
double les_constant(int model) {
  const double Csmag = 0.17, CWALE = 0.50;
  if (model == 11) return 2.0 * Csmag * Csmag * std::sqrt(2.0) * 9.0;
  if (model == 14) return CWALE * CWALE;
  return 2.5 * Csmag * Csmag;
}

// Asynchronous flow-file writers (replace the reference's fork per write,
// FluidDomain.f90:1702-1736): each write copies its packed fields into pinned
// host memory on the block's stream and a host thread writes the file once the
// copy's event has completed; fsil_write_wait joins them (the reference's
// mywait loop, FluidDomain.f90:355-357).
struct FlowWrite {
  std::thread th;
  explicit FlowWrite(std::thread &&t) : th(std::move(t)) {}
  FlowWrite(FlowWrite &&) = default;
  FlowWrite &operator=(FlowWrite &&) = default;
  ~FlowWrite() {
    if (th.joinable()) th.join();
  }
};
std::mutex g_write_mu;
std::vector<FlowWrite> g_writes;
std::string g_write_err;

void flow_write_job(hipEvent_t ev, char *host, size_t bytes, std::string path) {
  std::string err;
  if (hipEventSynchronize(ev) != hipSuccess) err = "flow write: device copy failed";
  if (err.empty()) {
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) err = "flow write: cannot open " + path;
    else {
      if (fwrite(host, 1, bytes, fp) != bytes) err = "flow write: short write to " + path;
      if (fclose(fp) != 0 && err.empty()) err = "flow write: cannot close " + path;
    }
  }
  (void)hipEventDestroy(ev);
  (void)hipHostFree(host);
  if (!err.empty()) {
    std::lock_guard<std::mutex> lk(g_write_mu);
    if (g_write_err.empty()) g_write_err = err;
  }
}

int plane_size(const Block &b, int face) {
  const int axis = face >> 1;
  return axis == 0 ? b.yd * b.zd : (axis == 1 ? b.xd * b.zd : b.xd * b.yd);
}

constexpr double kPi = 3.141592653589793;  // ConstParams.f90:36

// update_volume_force_ (FluidDomain.f90:1174-1180), host libm sin
void update_vf(Block &b) {
  b.VF[0] = b.VFin[0] + b.VFamp * std::sin(2.0 * kPi * b.VFfreq * b.blktime + b.VFphi / 180.0 * kPi);
  b.VF[1] = b.VFin[1];
  b.VF[2] = b.VFin[2];
}

// evaluate_oscillatory_velocity (FluidDomain.f90:1814-1824), host libm cos
void osc_velocity(const Block &b, double t, double v[3]) {
  v[0] = b.uvwIn[0] + b.shear[0] * std::cos(2 * kPi * b.shear[1] * t + b.shear[2] / 180.0 * kPi);
  v[1] = b.uvwIn[1];
  v[2] = b.uvwIn[2];
}

void eval_velocity_host(const Block &b, double t, double xc, double yc, double zc, double v[3]) {
  if (b.vkind == 0) {
    v[0] = ((b.uvwIn[0] + 0.0 * b.shear[0]) + yc * b.shear[1]) + zc * b.shear[2];
    v[1] = ((b.uvwIn[1] + xc * b.shear[0]) + 0.0 * b.shear[1]) + zc * b.shear[2];
    v[2] = ((b.uvwIn[2] + xc * b.shear[0]) + yc * b.shear[1]) + 0.0 * b.shear[2];
  } else if (b.vkind == 2) {
    osc_velocity(b, t, v);
  } else {
    v[0] = v[1] = v[2] = 0.0;
  }
}

bool is_active_face(int code) {
  return !(code == BC_PERIODIC || code == BC_FLUID || code == BC_FLUID_FATHER);
}

int launch_faces(Block &b) {
  auto args = [&](int face) {
    FaceArgs a{};
    const int code = b.bc[face];
    a.f = b.fA.p;
    a.g = b.gA.p;
    a.xd = b.xd; a.yd = b.yd; a.zd = b.zd; a.N = b.N;
    a.face = face; a.code = code;
    a.dh = b.dh; a.xmin = b.xmin; a.ymin = b.ymin; a.zmin = b.zmin;
    const double lo[3] = {b.xmin, b.ymin, b.zmin}, hi[3] = {b.xmax, b.ymax, b.zmax};
    const int axis = face >> 1;
    a.wallc = (face & 1) ? hi[axis] : lo[axis];
    a.hwc = (face & 1) ? hi[axis] + b.dh * 0.5 : lo[axis] - b.dh * 0.5;
    a.denIn = b.denIn;
    for (int k = 0; k < 3; ++k) { a.uvw[k] = b.uvwIn[k]; a.shear[k] = b.shear[k]; }
    a.vkind = b.vkind;
    if (b.vkind == 2) osc_velocity(b, b.blktime, a.uvw_osc);
    if (b.vkind != 0 && b.vkind != 2) a.vkind = 3;  // velocity left at zero
    a.nb_den = b.nb_den[face].p;
    a.nb_u = b.nb_u[face].p;
    a.hw = b.hw_valid[face] ? b.hw[face].p : nullptr;
    return a;
  };
  const dim3 block(256);
  const bool f32 = b.precision != 0;
  for (int axis = 0; axis < 3; ++axis) {
    const int lo = 2 * axis, hi = lo + 1;
    const bool alo = is_active_face(b.bc[lo]), ahi = is_active_face(b.bc[hi]);
    const int nda = axis == 0 ? b.xd : (axis == 1 ? b.yd : b.zd);
    const int nblk = (plane_size(b, lo) + 255) / 256;
    if (alo && ahi && nda >= 6) {
      // both faces of the axis in one launch (k_face_pair: disjoint planes)
      const FaceArgs A = args(lo), B = args(hi);
      const dim3 grid(2 * nblk);
      switch (axis + (f32 ? 3 : 0)) {
        case 0: hipLaunchKernelGGL((k_face_pair<0, false>), grid, block, 0, b.stream, A, B, nblk); break;
        case 1: hipLaunchKernelGGL((k_face_pair<1, false>), grid, block, 0, b.stream, A, B, nblk); break;
        case 2: hipLaunchKernelGGL((k_face_pair<2, false>), grid, block, 0, b.stream, A, B, nblk); break;
        case 3: hipLaunchKernelGGL((k_face_pair<0, true>), grid, block, 0, b.stream, A, B, nblk); break;
        case 4: hipLaunchKernelGGL((k_face_pair<1, true>), grid, block, 0, b.stream, A, B, nblk); break;
        case 5: hipLaunchKernelGGL((k_face_pair<2, true>), grid, block, 0, b.stream, A, B, nblk); break;
      }
      HIPCHK(hipGetLastError());
      continue;
    }
    for (int face = lo; face <= hi; ++face) {
      if (!is_active_face(b.bc[face])) continue;
      const FaceArgs a = args(face);
      const dim3 grid((plane_size(b, face) + 255) / 256);
      switch (face + (f32 ? 6 : 0)) {
        case 0: hipLaunchKernelGGL((k_face<0, false>), grid, block, 0, b.stream, a); break;
        case 1: hipLaunchKernelGGL((k_face<1, false>), grid, block, 0, b.stream, a); break;
        case 2: hipLaunchKernelGGL((k_face<2, false>), grid, block, 0, b.stream, a); break;
        case 3: hipLaunchKernelGGL((k_face<3, false>), grid, block, 0, b.stream, a); break;
        case 4: hipLaunchKernelGGL((k_face<4, false>), grid, block, 0, b.stream, a); break;
        case 5: hipLaunchKernelGGL((k_face<5, false>), grid, block, 0, b.stream, a); break;
        case 6: hipLaunchKernelGGL((k_face<0, true>), grid, block, 0, b.stream, a); break;
        case 7: hipLaunchKernelGGL((k_face<1, true>), grid, block, 0, b.stream, a); break;
        case 8: hipLaunchKernelGGL((k_face<2, true>), grid, block, 0, b.stream, a); break;
        case 9: hipLaunchKernelGGL((k_face<3, true>), grid, block, 0, b.stream, a); break;
        case 10: hipLaunchKernelGGL((k_face<4, true>), grid, block, 0, b.stream, a); break;
        case 11: hipLaunchKernelGGL((k_face<5, true>), grid, block, 0, b.stream, a); break;
      }
      HIPCHK(hipGetLastError());
    }
  }
  return 0;
}

dim3 cell_block(const Block &b) {
  // one block spans whole z-rows where possible so the z-shifted stores of a
  // row stay inside one workgroup (and one XCD's L2); longer rows are cut into
  // the width of 256, 192 or 128 that leaves the fewest idle lanes
  // (384-cell rows: two blocks of 192 rather than 256 + a half-idle 256)
  int tz = b.zd;
  if (b.zd > 256) {
    int best = 0;
    for (int w : {256, 192, 128}) {
      const int lanes = (b.zd + w - 1) / w * w;
      if (best == 0 || lanes < (b.zd + best - 1) / best * best) best = w;
    }
    tz = best;
  }
  int ty = std::max(1, 256 / tz);
  ty = std::min(ty, b.yd);
  return dim3(tz, ty, 1);
}

// Workgroup of the FP32-deviation kernel (at most 512 threads, ~70 VGPRs):
// rows of whole waves up to 512 cells are one workgroup (384-cell rows: six
// waves, so the row exchange applies), else the shape of cell_block.
dim3 cell_block32(const Block &b) {
  if (b.zd > 256 && b.zd <= 512 && b.zd % 64 == 0) return dim3(b.zd, 1, 1);
  return cell_block(b);
}

// The FP32-deviation kernel exchanges the z-moving values of a row through LDS when a
// workgroup is whole z-rows and the grid has no idle threads (every thread
// reaches the exchange's barrier); FSIL_ZROW=0 turns it off.  Resolved once at
// block creation, so the launch and the cache policy agree.
int zrow_exchange(const Block &b, dim3 blk) {
  const char *e = getenv("FSIL_ZROW");
  if (e && !strcmp(e, "0")) return 0;
  if (!((int)blk.x == b.zd && b.yd % (int)blk.y == 0 && blk.x * blk.y <= 512)) return 0;
  // rows of whole waves: lane permutes, 320 bytes of LDS (FSIL_ZROW=1 forces the LDS form)
  if (b.zd % 64 == 0 && !(e && !strcmp(e, "1"))) return 2;
  return 1;
}

// Cache policy of the fused kernel's population streams.  Measured on MI355X
// (both storage forms, periodic blocks): plain loads and stores are 8-12%
// faster at 384x192x192 and 512x256x256 (1-5 GB per population copy), where
// the caches merge the partial lines of the z-shifted stores; nontemporal ones
// are 3-14% faster at 1024x384x384 (11.5-23 GB per copy).  The switch is at
// 8 GiB per copy.  FP32-deviation blocks whose workgroups are whole z-rows
// store whole lines (zrow_exchange) and stream nontemporally at every size
// (384x192x192: 5.49 TB/s against 5.22 plain; 512x256x256: 5.76 against
// 5.43).  Blocks with coupled bodies always stream nontemporally
// (stream_hint_for): there the beam solver runs beside the collide kernel and
// the plain form's cache footprint slows the whole step (flexible plate,
// 384x192x192: 0.92 -> 1.19 ms FP32-deviation, 1.24 -> 1.43 ms FP64, though
// the collide kernel itself is 11% faster).  FSIL_STREAM_HINT=nt|plain
// overrides the choice (at block creation) for tests that run both.
int stream_policy(long long N, int precision, bool zrow) {
  const char *e = getenv("FSIL_STREAM_HINT");
  if (e && !strcmp(e, "nt")) return 1;
  if (e && !strcmp(e, "plain")) return 0;
  if (precision == 1 && zrow) return 1;  // whole-line stores: nontemporal is faster at every size measured
  const double copy_bytes = (double)Q * (double)N * (precision == 0 ? 8.0 : 4.0);
  return copy_bytes > 8.0 * 1024 * 1024 * 1024 ? 1 : 0;
}

// (Coupled launches once streamed nontemporally to spare the beam kernel's
// element data in the caches; that data now stays in the beam kernel's LDS,
// and the block's own policy is measured faster: FP64 flexible plate 1.187 ->
// 1.152 ms per step.)
int stream_hint_for(const Block &b, bool band) {
  (void)band;
  return b.nt;
}

template <int MODEL>
void launch_collide_model(const CollideArgs &a, bool force, bool band, bool aux, dim3 grid,
                          dim3 blk, hipStream_t s) {
#define L(F, B, X)                                                                                  \
  if (force == F && band == B && aux == X) {                                                        \
    if (a.nt) hipLaunchKernelGGL((k_collide_push<MODEL, F, B, X, true>), grid, blk, 0, s, a);        \
    else hipLaunchKernelGGL((k_collide_push<MODEL, F, B, X, false>), grid, blk, 0, s, a);            \
    return;                                                                                         \
  }
  L(false, false, false) L(true, false, false) L(false, true, false) L(true, true, false)
  L(false, false, true) L(true, false, true) L(false, true, true) L(true, true, true)
#undef L
}

// hull launch grid: the hole's row tiles (at least 256 workgroups); the
// kernel strides over a larger hull
unsigned hull_grid(const Block &b, const int *hole, dim3 blk) {
  const long long nz = (b.zd + blk.x - 1) / blk.x;
  const long long ny = (hole[3] - hole[2] + (int)blk.y) / (int)blk.y;
  const long long t = (long long)(hole[1] - hole[0] + 1) * ny * nz;
  return (unsigned)std::min<long long>(65535, std::max<long long>(256, t));
}

// the fused collide + push-stream launch of the whole block (region kRegionAll),
// of the cells outside `hole` (kRegionBulk; no band overlay, no swap), or of
// the hull of `hole` and the band box (kRegionHull)
int collide_stream(Block &b, int region = kRegionAll, const int *hole = nullptr) {
  CollideArgs a{};
  a.fin = b.fA.p;
  a.fout = b.fB.p;
  a.xd = b.xd; a.yd = b.yd; a.zd = b.zd; a.N = b.N;
  a.Omega = b.Omega; a.Omega2 = b.Omega2; a.dh = b.dh;
  a.c1 = 1.0 - 0.5 * b.Omega;
  a.hO = 0.5 * b.Omega; a.cT1 = 0.5 - 0.25 * b.Omega;
  a.hO2 = 0.5 * b.Omega2; a.cT2 = 0.5 - 0.25 * b.Omega2;
  for (int k = 0; k < 3; ++k) {
    a.hv[k] = 0.5 * b.VF[k] * b.dh;
    a.VF[k] = b.VF[k];
  }
  a.Mc = b.Mc.p;
  a.Mf = b.Mf.p;
  const bool force = (b.VF[0] != 0.0 || b.VF[1] != 0.0 || b.VF[2] != 0.0);
  const bool band = region != kRegionBulk && b.ibm.active && b.ibm.M > 0;
  a.nt = stream_hint_for(b, band || region != kRegionAll);
  if (band) {
    a.ctl = b.ibm.ctl.p;
    a.cellmap = b.ibm.cellmap.p;
    a.band_u = b.ibm.band_u.p;
    a.band_F = b.ibm.band_F.p;
    a.cap = b.ibm.cap;
  }
  a.region = region;
  if (region != kRegionAll) {
    if (!hole || b.ibm.M == 0 || is_les(b.model)) return fail("collide region %d without a band", region);
    for (int k = 0; k < 4; ++k) a.hole[k] = hole[k];
    a.ctl = b.ibm.ctl.p;
  }
  int m102 = 0, mhw = 0;
  for (int face = 0; face < 6; ++face) {
    if (b.bc[face] == BC_NEQ_DIRICHLET) m102 |= 1 << face;
    if (b.bc[face] == BC_WALL_HALFWAY || b.bc[face] == BC_MOVING_WALL_HALFWAY) mhw |= 1 << face;
    a.nb_den[face] = b.nb_den[face].p;
    a.nb_u[face] = b.nb_u[face].p;
    a.hw[face] = b.hw[face].p;
  }
  a.aux_face_mask102 = m102;
  a.aux_face_mask_hw = mhw;
  const bool aux = (m102 | mhw) != 0;
  const dim3 blk = cell_block(b);
  const dim3 grid = region == kRegionHull ? dim3(hull_grid(b, hole, blk))
                                          : dim3((b.zd + blk.x - 1) / blk.x, (b.yd + blk.y - 1) / blk.y, b.xd);
  if (is_les(b.model)) {
    a.tau_all = b.tau_all.p;
    a.uf = b.uf.p;
    a.tau = b.tau;
    a.nu = b.nu;
    a.les_c = les_constant(b.model);
    if (b.model != 11) {
      if (std::isnan(b.nu)) return fail("collision models 14 and 15 need flow%%nu (fsil_set_viscosity)");
      // the velocity field the gradients read, before any cell is collided
      if (band) hipLaunchKernelGGL(k_les_velocity<true>, grid, blk, 0, b.stream, a);
      else hipLaunchKernelGGL(k_les_velocity<false>, grid, blk, 0, b.stream, a);
      HIPCHK(hipGetLastError());
    }
  }
  if (b.model == 1) launch_collide_model<1>(a, force, band, aux, grid, blk, b.stream);
  else if (b.model == 2) launch_collide_model<2>(a, force, band, aux, grid, blk, b.stream);
  else if (b.model == 3) launch_collide_model<3>(a, force, band, aux, grid, blk, b.stream);
  else if (b.model == 11) launch_collide_model<11>(a, force, band, aux, grid, blk, b.stream);
  else if (b.model == 14) launch_collide_model<14>(a, force, band, aux, grid, blk, b.stream);
  else launch_collide_model<15>(a, force, band, aux, grid, blk, b.stream);
  HIPCHK(hipGetLastError());
  if (region == kRegionBulk) return 0;  // the hull launch completes the step
  for (int face = 0; face < 6; ++face)
    if ((mhw >> face) & 1) b.hw_valid[face] = true;
  std::swap(b.fA.p, b.fB.p);
  std::swap(b.fA.n, b.fB.n);
  return 0;
}

// FP32-deviation form; regions as collide_stream
int collide_stream32(Block &b, int region = kRegionAll, const int *hole = nullptr) {
  Collide32Args a{};
  a.gin = b.gA.p;
  a.gout = b.gB.p;
  a.xd = b.xd; a.yd = b.yd; a.zd = b.zd; a.N = b.N;
  a.Omega = (float)b.Omega;
  a.c1 = (float)(1.0 - 0.5 * b.Omega);
  a.hO = (float)(0.5 * b.Omega);
  a.hO2 = (float)(0.5 * b.Omega2);
  a.cT1 = (float)(0.5 - 0.25 * b.Omega);
  a.cT2 = (float)(0.5 - 0.25 * b.Omega2);
  a.Mc = b.Mc32.p;
  a.Mf = b.Mf32.p;
  a.dt3 = (float)(3.0 * b.dh);
  for (int k = 0; k < 3; ++k) {
    a.hv[k] = (float)(0.5 * b.VF[k] * b.dh);
    a.VF[k] = (float)b.VF[k];
  }
  const bool force = (b.VF[0] != 0.0 || b.VF[1] != 0.0 || b.VF[2] != 0.0);
  const bool band = region != kRegionBulk && b.ibm.active && b.ibm.M > 0;
  a.nt = stream_hint_for(b, band || region != kRegionAll);
  if (band) {
    a.ctl = b.ibm.ctl.p;
    a.cellmap = b.ibm.cellmap.p;
    a.band_u = b.ibm.band_u.p;
    a.band_F = b.ibm.band_F.p;
    a.cap = b.ibm.cap;
  }
  a.region = region;
  if (region != kRegionAll) {
    if (!hole || b.ibm.M == 0) return fail("collide region %d without a band", region);
    for (int k = 0; k < 4; ++k) a.hole[k] = hole[k];
    a.ctl = b.ibm.ctl.p;
  }
  int m102 = 0, mhw = 0;
  for (int face = 0; face < 6; ++face) {
    if (b.bc[face] == BC_NEQ_DIRICHLET) m102 |= 1 << face;
    if (b.bc[face] == BC_WALL_HALFWAY || b.bc[face] == BC_MOVING_WALL_HALFWAY) mhw |= 1 << face;
    a.nb_den[face] = b.nb_den[face].p;
    a.nb_u[face] = b.nb_u[face].p;
    a.hw[face] = b.hw[face].p;
  }
  a.aux_face_mask102 = m102;
  a.aux_face_mask_hw = mhw;
  const dim3 blk = cell_block32(b);
  const dim3 grid = region == kRegionHull ? dim3(hull_grid(b, hole, blk))
                                          : dim3((b.zd + blk.x - 1) / blk.x, (b.yd + blk.y - 1) / blk.y, b.xd);
  a.zrow = b.zrow;
  size_t lds = zrow32_lds_bytes(a.zrow, blk.x * blk.y);
  if (region == kRegionBulk) lds = std::max(lds, b.bulk_lds);  // (the kernel reads only the first bytes)
#define L32M(M, F, B)                                                                                  \
  if (a.nt) hipLaunchKernelGGL((k_collide_push32<M, F, B, true>), grid, blk, lds, b.stream, a);          \
  else hipLaunchKernelGGL((k_collide_push32<M, F, B, false>), grid, blk, lds, b.stream, a);
#define L32(F, B)                                                                                      \
  if (b.model == 2) { L32M(2, F, B) }                                                                   \
  else if (b.model == 3) { L32M(3, F, B) }                                                              \
  else { L32M(1, F, B) }
  if (force && band) { L32(true, true) }
  else if (force) { L32(true, false) }
  else if (band) { L32(false, true) }
  else { L32(false, false) }
#undef L32
#undef L32M
  HIPCHK(hipGetLastError());
  if (region == kRegionBulk) return 0;  // the hull launch completes the step
  for (int face = 0; face < 6; ++face)
    if ((mhw >> face) & 1) b.hw_valid[face] = true;
  std::swap(b.gA.p, b.gB.p);
  std::swap(b.gA.n, b.gB.n);
  return 0;
}

int macro32(Block &b, double *den, double *uuu) {
  DevBuf<double> dd, du;
  HIPCHK(dd.ensure(b.N));
  HIPCHK(du.ensure(3 * (size_t)b.N));
  hipLaunchKernelGGL(k_macro32, dim3((unsigned)((b.N + 255) / 256)), dim3(256), 0, b.stream, b.gA.p,
                     b.N, (float)(0.5 * b.VF[0] * b.dh), (float)(0.5 * b.VF[1] * b.dh),
                     (float)(0.5 * b.VF[2] * b.dh), dd.p, du.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(den, dd.p, b.N * sizeof(double), hipMemcpyDeviceToHost, b.stream));
  HIPCHK(hipMemcpyAsync(uuu, du.p, 3 * b.N * sizeof(double), hipMemcpyDeviceToHost, b.stream));
  HIPCHK(hipStreamSynchronize(b.stream));
  return 0;
}

// Peskin 4-point kernel (Solidbody.f90:822-833)
template <typename T>
int upload(DevBuf<T> &d, const T *h, size_t n, hipStream_t s) {
  HIPCHK(d.ensure(n));
  if (n) HIPCHK(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  return 0;
}

constexpr double kMPi = 3.141592653589793;  // m_pi (SolidSolver.f90:1261)

// Beam_Initialise + calculate_angle_material (isKB = 1) + InitTriadANDFormMass
// for one body, on the host (SolidSolver.f90:60-117, 1401-1510, 1533-1616)
void fem_init_body(const double *xyz8, int nND, int nEL, const int *elems, const int *cons, int iBodyModel,
                   const int *isMotionGiven, const double *mat, const double *kin, double Lref, double Uref,
                   double denIn, FemBodyHost &bh, std::vector<BeamElem> &els, std::vector<double> &pos,
                   std::vector<double> &vel) {
  const double denR = mat[0], KB = mat[1], KS = mat[2], psR = mat[3];
  bh.iBodyModel = iBodyModel;
  bh.Freq = kin[0];
  for (int k = 0; k < 3; ++k) {
    bh.XYZo[k] = kin[1 + k];
    bh.XYZAmpl[k] = kin[4 + k];
    bh.XYZPhi[k] = kin[7 + k] / 180.0 * kMPi;
    bh.AoAo[k] = kin[10 + k] / 180.0 * kMPi;
    bh.AoAAmpl[k] = kin[13 + k] / 180.0 * kMPi;
    bh.AoAPhi[k] = kin[16 + k] / 180.0 * kMPi;
    bh.initXYZVel[k] = kin[19 + k];
  }
  els.assign(nEL, BeamElem{});
  for (int n = 0; n < nEL; ++n) {
    BeamElem &e = els[n];
    const int p0 = elems[3 * n] - 1, p1 = elems[3 * n + 1] - 1;
    e.node0 = p0; e.node1 = p1; e.Nspan = elems[3 * n + 2];
    for (int k = 0; k < 3; ++k) {
      e.x00[k] = xyz8[8 * p0 + k];
      e.x00[6 + k] = xyz8[8 * p1 + k];
      e.x00[3 + k] = 0.0;
      e.x00[9 + k] = 0.0;
    }
    e.area00 = std::sqrt(bpow(e.x00[6] - e.x00[0], 2) + bpow(e.x00[7] - e.x00[1], 2) + bpow(e.x00[8] - e.x00[2], 2));
    for (int k = 0; k < 6; ++k) { e.bc[k] = cons[6 * p0 + k]; e.bc[6 + k] = cons[6 * p1 + k]; }
    e.Lspan = 0.5 * (xyz8[8 * p0 + 3] + xyz8[8 * p1 + 3]);
    e.spanlen = 0.5 * (xyz8[8 * p0 + 4] + xyz8[8 * p1 + 4]) + e.Lspan;
    for (int k = 0; k < 3; ++k) e.dirc00[k] = 0.5 * (xyz8[8 * p0 + 5 + k] + xyz8[8 * p1 + 5 + k]);
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + bpow(e.dirc00[k], 2);
    const double nrm = std::sqrt(s);
    for (int k = 0; k < 3; ++k) e.dirc00[k] = e.dirc00[k] / nrm;
    if (e.bc[0] == 1) for (int k = 0; k < 6; ++k) e.bc[k] = isMotionGiven[k];
    if (e.bc[6] == 1) for (int k = 0; k < 6; ++k) e.bc[6 + k] = isMotionGiven[k];
    const double len = e.spanlen;
    const double th = std::sqrt(KB / KS * 12.0) * Lref;
    e.prop[2] = len * th;
    e.prop[0] = KS * denIn * bpow(Uref, 2) * Lref * len / e.prop[2];
    e.prop[1] = e.prop[0] / (2.0 * (1.0 + psR));
    e.prop[3] = denR * len * Lref * denIn / e.prop[2];
    const double ratio = th / len;
    e.prop[5] = len * bpow(th, 3) / 3.0 * (1.0 - 0.63 * ratio + 0.052 * bpow(ratio, 5));
    e.prop[6] = len * bpow(th, 3) / 12.0;
    e.prop[7] = th * bpow(len, 3) / 12.0;
  }
  const double t = 0.0;
  double XYZ[3], AoA[3], AoAd[3], Tn[9];
  for (int k = 0; k < 3; ++k) {
    XYZ[k] = bh.XYZo[k] + bh.XYZAmpl[k] * std::cos(2.0 * kMPi * bh.Freq * t + bh.XYZPhi[k]) + bh.initXYZVel[k] * t;
    AoA[k] = bh.AoAo[k] + bh.AoAAmpl[k] * std::cos(2.0 * kMPi * bh.Freq * t + bh.AoAPhi[k]);
  }
  aoa_to_ttt(AoA, bh.TTT0);
  aoa_to_ttt(AoA, Tn);
  angle_triad(bh.TTT0, Tn, AoAd);
  for (auto &e : els) {
    el_map_pos(e, e.x0, bh.TTT0, XYZ, AoAd);
    bmv3(bh.TTT0, e.dirc00, e.dirc0);
    for (int k = 0; k < NED; ++k) { e.x1[k] = e.x0[k]; e.xnxt[k] = e.x1[k]; }
    e.dx0 = e.x0[6] - e.x0[0];
    e.dy0 = e.x0[7] - e.x0[1];
    e.dz0 = e.x0[8] - e.x0[2];
    e.len0 = std::sqrt(e.dx0 * e.dx0 + e.dy0 * e.dy0 + e.dz0 * e.dz0);
    e.xll0 = e.dx0 / e.len0;
    e.xmm0 = e.dy0 / e.len0;
    e.xnn0 = e.dz0 / e.len0;
    el_cptdxyz1(e);
  }
  pos.assign(6 * (size_t)nND, 0.0);
  vel.assign(6 * (size_t)nND, 0.0);
  for (auto &e : els)
    for (int k = 0; k < 6; ++k) { pos[6 * e.node0 + k] = e.x1[k]; pos[6 * e.node1 + k] = e.x1[6 + k]; }
  if (iBodyModel == 1) {
    double UVW[3], W1[3], W2[3], W3[3];
    for (int k = 0; k < 3; ++k) {
      UVW[k] = -(2.0 * kMPi * bh.Freq * bh.XYZAmpl[k] * std::sin(2.0 * kMPi * bh.Freq * t + bh.XYZPhi[k])) + bh.initXYZVel[k];
      W1[k] = -(2.0 * kMPi * bh.Freq * bh.AoAAmpl[k] * std::sin(2.0 * kMPi * bh.Freq * t + bh.AoAPhi[k]));
    }
    W2[0] = W1[0] * std::cos(AoA[1]) + W1[2];
    W2[1] = W1[0] * std::sin(AoA[1]) * std::sin(AoA[2]) + W1[1] * std::cos(AoA[2]);
    W2[2] = W1[0] * std::sin(AoA[1]) * std::cos(AoA[2]) - W1[1] * std::sin(AoA[2]);
    bmv3(bh.TTT0, W2, W3);
    for (int n = 0; n < nND; ++n) {
      double rel[3];
      for (int k = 0; k < 3; ++k) rel[k] = pos[6 * n + k] - XYZ[k];
      vel[6 * n + 0] = (W3[1] * rel[2] - W3[2] * rel[1]) + UVW[0];
      vel[6 * n + 1] = (W3[2] * rel[0] - W3[0] * rel[2]) + UVW[1];
      vel[6 * n + 2] = (W3[0] * rel[1] - W3[1] * rel[0]) + UVW[2];
      for (int k = 0; k < 3; ++k) vel[6 * n + 3 + k] = W3[k];
    }
  }
  for (auto &e : els) {
    axis_dir_triad(e.xll0, e.xmm0, e.xnn0, e.dirc0, e.tn1);
    for (int k = 0; k < 9; ++k) { e.tn2[k] = e.tn1[k]; e.tee[k] = e.tn1[k]; }
    el_form_mass(e);
    el_rotate(e);
    el_rkr(e, e.mas);
    el_form_stiff(e);
    for (int k = 0; k < NED * NED; ++k) e.stl[k] = e.stf[k];
    (void)el_stiff_scalars(e);  // checked by the caller
  }
}

// prescribed placement of a sub-step, host libm (SolidSolver.f90:1826-1869)
void fem_kinematics(const FemBodyHost &bh, double time, int isub, double deltat, double subdt, FemBodyDesc &d) {
  const double tt = time - deltat + (double)isub * subdt;
  double AoA[3];
  for (int k = 0; k < 3; ++k) {
    d.XYZ[k] = bh.XYZo[k] + bh.XYZAmpl[k] * std::cos(2.0 * kMPi * bh.Freq * tt + bh.XYZPhi[k]) + bh.initXYZVel[k] * tt;
    AoA[k] = bh.AoAo[k] + bh.AoAAmpl[k] * std::cos(2.0 * kMPi * bh.Freq * tt + bh.AoAPhi[k]);
  }
  aoa_to_ttt(AoA, d.T);
  angle_triad(bh.TTT0, d.T, d.AoAd);
  if (bh.iBodyModel == 1) {
    double W1[3], W2[3];
    for (int k = 0; k < 3; ++k) {
      d.UVW[k] = -(2.0 * kMPi * bh.Freq * bh.XYZAmpl[k] * std::sin(2.0 * kMPi * bh.Freq * tt + bh.XYZPhi[k])) + bh.initXYZVel[k];
      W1[k] = -(2.0 * kMPi * bh.Freq * bh.AoAAmpl[k] * std::sin(2.0 * kMPi * bh.Freq * tt + bh.AoAPhi[k]));
    }
    W2[0] = W1[0] * std::cos(AoA[1]) + W1[2];
    W2[1] = W1[0] * std::sin(AoA[1]) * std::sin(AoA[2]) + W1[1] * std::cos(AoA[2]);
    W2[2] = W1[0] * std::sin(AoA[1]) * std::cos(AoA[2]) - W1[1] * std::sin(AoA[2]);
    bmv3(d.T, W2, d.W3);
  }
}

FemArrays fem_arrays(Block &b) {
  Fem &F = b.fem;
  FemArrays A{};
  A.el = F.el.p;
  A.pos = F.pos.p; A.dsp = F.dsp.p; A.vel = F.vel.p; A.acc = F.acc.p; A.lod = b.ibm.lod.p;
  A.dspO = F.dspO.p; A.velO = F.velO.p; A.accO = F.accO.p; A.lodInte = F.lodInte.p; A.lodEffe = F.lodEffe.p;
  A.lodExte = F.lodExte.p; A.vBC = F.vBC.p; A.x = F.x.p; A.r = F.r.p; A.p = F.p.p; A.Ap = F.Ap.p; A.z = F.z.p;
  A.xF = F.xF.p; A.lbuf = F.lbuf.p; A.blk = F.blk.p; A.fixed = F.fixed.p;
  A.dof_ptr = F.dof_ptr.p; A.dof_e = F.dof_e.p; A.dof_i = F.dof_i.p;
  A.node_ptr = F.node_ptr.p; A.node_e = F.node_e.p; A.node_h = F.node_h.p;
  return A;
}

int fem_collect(Block &b);
int fem_order_after(Block &b);
int fill_solid_info_args(Block &b, int body, const double *refs, double *buf, SolidInfoArgs &a);
int ibm_check_report(Block &b);

}  // namespace

// ===========================================================================
extern "C" {

const char *fsil_last_error(void) { return g_err.c_str(); }

int fsil_version(void) { return 1; }

int fsil_init(int device) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (n <= 0) return fail("no HIP device visible");
  if (device < 0 || device >= n) return fail("device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  g_device = device;
  if (!g_scratch.count(device)) {
    double *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail("no device address for the static scratch on device %d", device);
    g_scratch[device] = ptr;
  }
  return 0;
}

int fsil_finalize(void) {
  (void)fsil_write_wait();
  for (auto &kv : g_blocks) {
    Block &b = *kv.second;
    if (b.stream) (void)hipStreamSynchronize(b.stream);
    if (b.ev0) (void)hipEventDestroy(b.ev0);
    if (b.ev1) (void)hipEventDestroy(b.ev1);
    if (b.ev_fluid) (void)hipEventDestroy(b.ev_fluid);
  }
  g_blocks.clear();
  return 0;
}

int fsil_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int fsil_block_create(int id, const int *dims, const double *geom, const int *bc, int iCollidModel,
                      const double *relax, const double *m_collid, const double *m_force,
                      int precision, int *handle) {
  if (g_device < 0) return fail("fsil_init has not been called");
  if (!dims || !geom || !bc || !relax || !handle) return fail("null argument to fsil_block_create");
  if (dims[0] < 4 || dims[1] < 4 || dims[2] < 4) return fail("block dimensions must be >= 4");
  if (dims[0] > 32767 || dims[1] > 32767 || dims[2] > 32767)
    return fail("Grid number exceeds 32767");  // FluidDomain.f90:88-91
  if (!(iCollidModel == 1 || iCollidModel == 2 || iCollidModel == 3 || is_les(iCollidModel)))
    return fail("collision model %d not provided (1 SRT, 2 TRT, 3 MRT, 11 Smagorinsky, 14 WALE, 15 Vreman)",
                iCollidModel);
  if (iCollidModel == 3 && (!m_collid || !m_force)) return fail("MRT needs M_COLLID and M_FORCE");
  if (!(precision == 0 || precision == 1)) return fail("precision must be 0 (FP64) or 1 (FP32 deviation)");
  if (precision == 1 && !(iCollidModel >= 1 && iCollidModel <= 3))
    return fail("FP32-deviation form supports collision models 1-3 (SRT, TRT, MRT), not %d", iCollidModel);
  if (precision == 1 && (long long)dims[0] * dims[1] * dims[2] > (1LL << 30))
    return fail("FP32-deviation blocks hold at most 2^30 cells (32-bit byte offsets per population plane)");
  if (precision == 0 && (long long)dims[0] * dims[1] * dims[2] > (1LL << 29))
    return fail("FP64 blocks hold at most 2^29 cells (32-bit byte offsets per population plane)");
  for (int k = 0; k < 6; ++k) {
    const int c = bc[k];
    const bool ok = c == 0 || c == 1 || c == 101 || c == 102 || c == 103 || c == 104 || c == 201 ||
                    c == 202 || c == 203 || c == 204 || c == 301 || c == 302;
    if (!ok) return fail("face %d has no such boundary condition %d", k + 1, c);
  }
  for (int a = 0; a < 3; ++a) {
    const bool p0 = bc[2 * a] == BC_PERIODIC, p1 = bc[2 * a + 1] == BC_PERIODIC;
    if (p0 != p1) return fail("Periodic boundaries must appear in pairs: %d %d", bc[2 * a], bc[2 * a + 1]);
  }
  auto bp = std::make_unique<Block>();
  Block &b = *bp;
  b.id = id;
  b.xd = dims[0]; b.yd = dims[1]; b.zd = dims[2];
  b.N = (long long)b.xd * b.yd * b.zd;
  b.dh = geom[0]; b.xmin = geom[1]; b.ymin = geom[2]; b.zmin = geom[3];
  b.xmax = geom[4]; b.ymax = geom[5]; b.zmax = geom[6];
  for (int k = 0; k < 6; ++k) b.bc[k] = bc[k];
  b.model = iCollidModel;
  b.tau = relax[0]; b.Omega = relax[1]; b.Omega2 = relax[2];
  b.precision = precision;
  b.zrow = precision == 1 ? zrow_exchange(b, cell_block32(b)) : 0;
  b.nt = stream_policy(b.N, precision, b.zrow != 0);
  b.nt_forced = getenv("FSIL_STREAM_HINT") != nullptr;
  b.sh = std::make_shared<StreamHolder>();
  HIPCHK(hipStreamCreateWithFlags(&b.sh->s, hipStreamNonBlocking));
  b.stream = b.sh->s;
  HIPCHK(hipEventCreate(&b.ev0));
  HIPCHK(hipEventCreate(&b.ev1));
  HIPCHK(hipEventCreateWithFlags(&b.ev_fluid, hipEventDisableTiming));
  {
    // Split steps pay where the coupling and beam chain is the step's critical
    // path (FP32-deviation: 0.781 against 0.807 ms); in FP64 the fluid update
    // is, and the split costs (flexible plate 1.174 against 1.152 ms):
    // FSIL_SPLIT=1 or 0 overrides.
    const char *e = getenv("FSIL_SPLIT");
    b.split_ok = e ? strcmp(e, "0") != 0 : precision == 1;
    const char *m = getenv("FSIL_SPLIT_MARGIN");
    if (m) b.split_margin = atoi(m);
    // The bulk launch of an FP32 split step runs beside the coupling call,
    // whose small launches wait for CU slots while the bulk's workgroups fill
    // every CU.  The bulk asks for enough LDS that at most R of its
    // workgroups share a CU (FSIL_BULK_RESIDENCY=R, default 2; 0: no limit).
    // Measured on MI355X, flexible plate 384x192x192: 0.714 ms per step
    // unlimited (6 per CU), 0.689 at 4, 0.685 at 3, 0.675 at 2; the bulk
    // itself takes 489 us at 2 against 405 us, the coupling's band build
    // 57 us against 200 us.
    int R = 2;
    if (const char *r = getenv("FSIL_BULK_RESIDENCY")) R = atoi(r);
    int dev = 0, per_cu = 0;
    HIPCHK(hipGetDevice(&dev));
    if (R > 0 && hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
                     hipSuccess && per_cu > 0) {
      const size_t want = (size_t)per_cu / (size_t)(R + 1) + 1024;
      if (want <= 64 * 1024) b.bulk_lds = want;  // within the default dynamic-LDS limit
    }
    (void)hipGetLastError();
  }
  const size_t nf = (size_t)Q * b.N;
  if (precision == 0) {
    HIPCHK(b.fA.ensure(nf));
    HIPCHK(b.fB.ensure(nf));
  } else {
    HIPCHK(b.gA.ensure(nf));
    HIPCHK(b.gB.ensure(nf));
  }
  if (is_les(iCollidModel)) {
    // tau_all = tau (calculate_SRT_params, FluidDomain.f90:450-456)
    HIPCHK(b.tau_all.ensure(b.N));
    hipLaunchKernelGGL(k_fill_value, dim3((unsigned)((b.N + 255) / 256)), dim3(256), 0, b.stream, b.tau_all.p, b.N,
                       b.tau);
    HIPCHK(hipGetLastError());
    if (iCollidModel != 11) HIPCHK(b.uf.ensure(3 * (size_t)b.N));
    HIPCHK(hipStreamSynchronize(b.stream));
  }
  if (iCollidModel == 3) {
    HIPCHK(b.Mc.ensure(Q * Q));
    HIPCHK(b.Mf.ensure(Q * Q));
    HIPCHK(hipMemcpy(b.Mc.p, m_collid, Q * Q * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.Mf.p, m_force, Q * Q * sizeof(double), hipMemcpyHostToDevice));
    if (precision == 1) {
      float mc[Q * Q], mf[Q * Q];
      for (int k = 0; k < Q * Q; ++k) {
        mc[k] = (float)m_collid[k];
        mf[k] = (float)m_force[k];
      }
      HIPCHK(b.Mc32.ensure(Q * Q));
      HIPCHK(b.Mf32.ensure(Q * Q));
      HIPCHK(hipMemcpy(b.Mc32.p, mc, sizeof(mc), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(b.Mf32.p, mf, sizeof(mf), hipMemcpyHostToDevice));
    }
  }
  for (int face = 0; face < 6; ++face) {
    const int np = plane_size(b, face);
    if (bc[face] == BC_NEQ_DIRICHLET) {
      HIPCHK(b.nb_den[face].ensure(np));
      HIPCHK(b.nb_u[face].ensure(3 * (size_t)np));
    }
    if (bc[face] == BC_WALL_HALFWAY || bc[face] == BC_MOVING_WALL_HALFWAY)
      HIPCHK(b.hw[face].ensure((size_t)Q * np));
  }
  const int h = g_next_handle++;
  g_blocks[h] = std::move(bp);
  *handle = h;
  return 0;
}

int fsil_block_destroy(int h) {
  Block *b = get_block(h);
  if (!b) return 1;
  (void)hipStreamSynchronize(b->stream);
  (void)hipEventDestroy(b->ev0);
  (void)hipEventDestroy(b->ev1);
  (void)hipEventDestroy(b->ev_fluid);
  // drop the pair that names this block as a son, and detach its own sons
  if (b->father) {
    auto it = g_blocks.find(b->father);
    if (it != g_blocks.end()) {
      auto &v = it->second->sons;
      v.erase(std::remove_if(v.begin(), v.end(), [h](const std::unique_ptr<Pair> &p) { return p->son == h; }),
              v.end());
    }
  }
  for (auto &p : b->sons) {
    auto it = g_blocks.find(p->son);
    if (it != g_blocks.end()) it->second->father = 0;
  }
  g_blocks.erase(h);
  return 0;
}

int fsil_set_flow(int h, double denIn, const double *uvwIn, const double *shearRateIn,
                  int velocityKind, const double *volumeForceIn, double volumeForceAmp,
                  double volumeForceFreq, double volumeForcePhi) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!uvwIn || !shearRateIn || !volumeForceIn) return fail("null argument to fsil_set_flow");
  b->denIn = denIn;
  for (int k = 0; k < 3; ++k) {
    b->uvwIn[k] = uvwIn[k];
    b->shear[k] = shearRateIn[k];
    b->VFin[k] = volumeForceIn[k];
  }
  b->vkind = velocityKind;
  b->VFamp = volumeForceAmp;
  b->VFfreq = volumeForceFreq;
  b->VFphi = volumeForcePhi;
  update_vf(*b);
  return 0;
}

int fsil_set_time(int h, double blktime) {
  Block *b = get_block(h);
  if (!b) return 1;
  b->blktime = blktime;
  update_vf(*b);
  return 0;
}

int fsil_set_viscosity(int h, double nu) {
  Block *b = get_block(h);
  if (!b) return 1;
  b->nu = nu;
  return 0;
}

int fsil_tau_download(int h, double *tau_all) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!tau_all) return fail("null tau_all");
  if (!is_les(b->model)) {
    for (long long n = 0; n < b->N; ++n) tau_all[n] = b->tau;
    return 0;
  }
  HIPCHK(hipMemcpyAsync(tau_all, b->tau_all.p, b->N * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_get_volume_force(int h, double *vf) {
  Block *b = get_block(h);
  if (!b) return 1;
  for (int k = 0; k < 3; ++k) vf[k] = b->VF[k];
  return 0;
}

int fsil_f_upload(int h, const double *fIn) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!fIn) return fail("null fIn");
  const size_t nf = (size_t)Q * b->N;
  if (b->precision == 0) {
    HIPCHK(hipMemcpyAsync(b->fA.p, fIn, nf * sizeof(double), hipMemcpyHostToDevice, b->stream));
  } else {
    // FP32 deviation storage: g = f - w, rounded once
    std::vector<float> g(nf);
    for (int q = 0; q < Q; ++q)
      for (long long n = 0; n < b->N; ++n) g[q * b->N + n] = (float)(fIn[q * b->N + n] - wt(q));
    HIPCHK(hipMemcpyAsync(b->gA.p, g.data(), nf * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_f_fill(int h, const double *f19) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!f19) return fail("null populations");
  const dim3 grid((unsigned)((b->N + 255) / 256));
  if (b->precision == 0) {
    Fill19 a;
    for (int q = 0; q < Q; ++q) a.v[q] = f19[q];
    hipLaunchKernelGGL(k_fill19, grid, dim3(256), 0, b->stream, b->fA.p, b->N, a);
  } else {
    Fill19f a;  // FP32 deviation storage: g = f - w, rounded once (as fsil_f_upload)
    for (int q = 0; q < Q; ++q) a.v[q] = (float)(f19[q] - wt(q));
    hipLaunchKernelGGL(k_fill19f, grid, dim3(256), 0, b->stream, b->gA.p, b->N, a);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_f_download(int h, double *fIn) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!fIn) return fail("null fIn");
  const size_t nf = (size_t)Q * b->N;
  if (b->precision == 0) {
    HIPCHK(hipMemcpyAsync(fIn, b->fA.p, nf * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  } else {
    std::vector<float> g(nf);
    HIPCHK(hipMemcpyAsync(g.data(), b->gA.p, nf * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    for (int q = 0; q < Q; ++q)
      for (long long n = 0; n < b->N; ++n) fIn[q * b->N + n] = (double)g[q * b->N + n] + wt(q);
  }
  return 0;
}

int fsil_set_boundary_conditions(int h) {
  Block *b = get_block(h);
  if (!b) return 1;
  // Code 102 reads den/uuu of the first interior plane.  At start-up the
  // reference calls this right after initialise_, where den = denIn and uuu is
  // the inflow profile at every node (FluidDomain.f90:533-545), so those fields
  // are what the planes hold here.
  for (int face = 0; face < 6; ++face) {
    if (b->bc[face] != BC_NEQ_DIRICHLET) continue;
    const int np = plane_size(*b, face);
    std::vector<double> den(np), uu(3 * (size_t)np);
    const int axis = face >> 1, hi = face & 1;
    const int nda = axis == 0 ? b->xd : (axis == 1 ? b->yd : b->zd);
    const int p2 = hi ? nda - 2 : 1;
    const int n0 = axis == 0 ? b->yd : b->xd, n1 = axis == 2 ? b->yd : b->zd;
    for (int A = 0; A < n0; ++A)
      for (int B = 0; B < n1; ++B) {
        int c[3];
        c[axis] = p2;
        c[axis == 0 ? 1 : 0] = A;
        c[axis == 2 ? 1 : 2] = B;
        double v[3];
        eval_velocity_host(*b, 0.0, b->xmin + b->dh * c[0], b->ymin + b->dh * c[1],
                           b->zmin + b->dh * c[2], v);
        const int t = A * n1 + B;
        den[t] = b->denIn;
        uu[t] = v[0]; uu[np + t] = v[1]; uu[2 * (size_t)np + t] = v[2];
      }
    HIPCHK(hipMemcpyAsync(b->nb_den[face].p, den.data(), np * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->nb_u[face].p, uu.data(), 3 * np * sizeof(double), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  if (launch_faces(*b)) return 1;
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_macro_download(int h, double *den, double *uuu) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!den || !uuu) return fail("null den/uuu");
  if (b->precision != 0) return macro32(*b, den, uuu);
  DevBuf<double> dd, du;
  HIPCHK(dd.ensure(b->N));
  HIPCHK(du.ensure(3 * (size_t)b->N));
  const double hv0 = 0.5 * b->VF[0] * b->dh, hv1 = 0.5 * b->VF[1] * b->dh, hv2 = 0.5 * b->VF[2] * b->dh;
  hipLaunchKernelGGL(k_macro, dim3((unsigned)((b->N + 255) / 256)), dim3(256), 0, b->stream, b->fA.p,
                     b->N, hv0, hv1, hv2, dd.p, du.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(den, dd.p, b->N * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(uuu, du.p, 3 * b->N * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

namespace {
// DatFlow fields of block b over its output window into the device staging
// buffer (on the block's stream); nw = cells in the window
int flow_pack(Block &b, int off, double denIn, double Uref, long long &nw) {
  if (off < 0 || 2 * off >= b.xd || 2 * off >= b.yd || 2 * off >= b.zd)
    return fail("offsetOutput %d leaves no cells of a %dx%dx%d block", off, b.xd, b.yd, b.zd);
  nw = (long long)(b.xd - 2 * off) * (b.yd - 2 * off) * (b.zd - 2 * off);
  HIPCHK(b.outbuf.ensure(4 * (size_t)nw));
  const dim3 grid((unsigned)((nw + 255) / 256)), blk(256);
  const double hv0 = 0.5 * b.VF[0] * b.dh, hv1 = 0.5 * b.VF[1] * b.dh, hv2 = 0.5 * b.VF[2] * b.dh;
  const double inv = 1.0 / Uref;
  if (b.precision == 0)
    hipLaunchKernelGGL(k_flow_pack<false>, grid, blk, 0, b.stream, b.fA.p, nullptr, b.N, b.xd, b.yd, b.zd, off, hv0,
                       hv1, hv2, denIn, inv, b.outbuf.p);
  else
    hipLaunchKernelGGL(k_flow_pack<true>, grid, blk, 0, b.stream, nullptr, b.gA.p, b.N, b.xd, b.yd, b.zd, off, hv0,
                       hv1, hv2, denIn, inv, b.outbuf.p);
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace

int fsil_flow_pack(int h, int offsetOutput, double denIn, double Uref, float *out) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!out) return fail("null output");
  long long nw = 0;
  if (flow_pack(*b, offsetOutput, denIn, Uref, nw)) return 1;
  HIPCHK(hipMemcpyAsync(out, b->outbuf.p, 4 * (size_t)nw * sizeof(float), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_write_flow(int h, int offsetOutput, double denIn, double Uref, const char *path) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (!path) return fail("null path");
  long long nw = 0;
  if (flow_pack(*b, offsetOutput, denIn, Uref, nw)) return 1;
  const size_t head = 4 * sizeof(int) + 4 * sizeof(double), data = 4 * (size_t)nw * sizeof(float);
  char *host = nullptr;
  HIPCHK(hipHostMalloc((void **)&host, head + data, hipHostMallocDefault));
  const int ih[4] = {b->xd - 2 * offsetOutput, b->yd - 2 * offsetOutput, b->zd - 2 * offsetOutput, b->id};
  const double dh[4] = {b->xmin + offsetOutput * b->dh, b->ymin + offsetOutput * b->dh,
                        b->zmin + offsetOutput * b->dh, b->dh};
  memcpy(host, ih, sizeof(ih));
  memcpy(host + sizeof(ih), dh, sizeof(dh));
  hipEvent_t ev = nullptr;
  hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMemcpyAsync(host + head, b->outbuf.p, data, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipEventRecord(ev, b->stream);
  if (e != hipSuccess) {
    if (ev) (void)hipEventDestroy(ev);
    (void)hipHostFree(host);
    return fail("fsil_write_flow: %s", hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(g_write_mu);
  g_writes.emplace_back(std::thread(flow_write_job, ev, host, head + data, std::string(path)));
  return 0;
}

int fsil_write_wait(void) {
  std::vector<FlowWrite> w;
  {
    std::lock_guard<std::mutex> lk(g_write_mu);
    w.swap(g_writes);
  }
  for (auto &x : w)
    if (x.th.joinable()) x.th.join();
  std::lock_guard<std::mutex> lk(g_write_mu);
  if (!g_write_err.empty()) {
    const std::string e = g_write_err;
    g_write_err.clear();
    return fail("%s", e.c_str());
  }
  return 0;
}

int fsil_write_continue(int nblocks, const int *handles, int step, double time, const char *path) {
  if (nblocks <= 0 || !handles || !path) return fail("bad arguments to fsil_write_continue");
  FILE *fp = fopen(path, "wb");
  if (!fp) return fail("cannot open %s", path);
  int rc = 0;
  const int hdr[2] = {nblocks, step};
  if (fwrite(hdr, sizeof(int), 2, fp) != 2 || fwrite(&time, sizeof(double), 1, fp) != 1) rc = fail("write error on %s", path);
  std::vector<double> f;
  for (int k = 0; k < nblocks && rc == 0; ++k) {
    Block *b = get_block(handles[k]);
    if (!b) { rc = 1; break; }
    const double g[4] = {b->xmin, b->ymin, b->zmin, b->dh};
    const int d[3] = {b->xd, b->yd, b->zd};
    f.resize((size_t)Q * b->N);
    if (fsil_f_download(handles[k], f.data())) { rc = 1; break; }
    if (fwrite(g, sizeof(double), 4, fp) != 4 || fwrite(d, sizeof(int), 3, fp) != 3 ||
        fwrite(f.data(), sizeof(double), f.size(), fp) != f.size())
      rc = fail("write error on %s", path);
  }
  if (fclose(fp) != 0 && rc == 0) rc = fail("cannot close %s", path);
  return rc;
}

int fsil_collide_stream_bc(int h) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (b->precision == 0) {
    if (collide_stream(*b)) return 1;
    if (launch_faces(*b)) return 1;
  } else {
    if (collide_stream32(*b)) return 1;
    if (launch_faces(*b)) return 1;
  }
  b->ibm.active = false;
  return 0;
}

int fsil_steps(int h, int nsteps, double t0, double dt) {
  Block *b = get_block(h);
  if (!b) return 1;
  double t = t0;
  for (int s = 0; s < nsteps; ++s) {
    t = t + dt;  // main.f90:95 accumulates the time
    b->blktime = t;
    update_vf(*b);
    if (fsil_collide_stream_bc(h)) return 1;
  }
  return 0;
}

int fsil_time_collide(int h, int nlaunch, double *avg_ms) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (nlaunch <= 0 || !avg_ms) return fail("bad arguments to fsil_time_collide");
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  for (int s = 0; s < nlaunch; ++s) {
    if (b->precision == 0) {
      if (collide_stream(*b)) return 1;
    } else {
      if (collide_stream32(*b)) return 1;
    }
  }
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  *avg_ms = (double)ms / nlaunch;
  return 0;
}

int fsil_time_steps(int h, int nsteps, double t0, double dt, double *total_ms) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (nsteps <= 0 || !total_ms) return fail("bad arguments to fsil_time_steps");
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  if (fsil_steps(h, nsteps, t0, dt)) return 1;
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  *total_ms = ms;
  return 0;
}

// ---------------------------------------------------------------------------
// immersed boundary
int fsil_ibm_define_bodies(int h, int nbodies, const int *nmarkers, const int *nnodes,
                           const int *nelems, const int *marker_elem, const int *elem_nodes) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (nbodies < 0) return fail("negative body count");
  Ibm &I = b->ibm;
  I = Ibm();
  I.nbodies = nbodies;
  // entries per LDS pass of the band-cell sums (tests force the per-cell form with a small value)
  I.pass_cap = kPassEntries;
  if (const char *e = getenv("FSIL_BAND_PASS")) I.pass_cap = std::max(0, std::min(kPassEntries, atoi(e)));
  I.nmark.assign(nmarkers, nmarkers + nbodies);
  I.nnode.assign(nnodes, nnodes + nbodies);
  I.nelem.assign(nelems, nelems + nbodies);
  I.moff.assign(nbodies + 1, 0);
  I.noff.assign(nbodies + 1, 0);
  I.eoff.assign(nbodies + 1, 0);
  for (int k = 0; k < nbodies; ++k) {
    I.moff[k + 1] = I.moff[k] + I.nmark[k];
    I.noff[k + 1] = I.noff[k] + I.nnode[k];
    I.eoff[k + 1] = I.eoff[k] + I.nelem[k];
  }
  I.M = I.moff[nbodies];
  I.nnodes = I.noff[nbodies];
  I.nelems = I.eoff[nbodies];
  if ((long long)I.M * 64 > (1LL << 30)) return fail("too many markers");
  // global element id per marker; DOF lists (lodFlow) in marker order
  I.marker_elem.resize(I.M);
  I.ndof = 6 * I.nnodes;
  std::vector<std::vector<int>> lists(I.ndof);
  for (int body = 0; body < nbodies; ++body)
    for (int m = I.moff[body]; m < I.moff[body + 1]; ++m) {
      const int el = marker_elem[m];
      if (el < 0 || el >= I.nelem[body]) return fail("marker %d has element %d out of range", m, el);
      const int ge = I.eoff[body] + el;
      I.marker_elem[m] = ge;
      const int n0 = elem_nodes[2 * ge], n1 = elem_nodes[2 * ge + 1];
      if (n0 < 0 || n0 >= I.nnode[body] || n1 < 0 || n1 >= I.nnode[body])
        return fail("element %d has a node out of range", el);
      const int g0 = 6 * (I.noff[body] + n0), g1 = 6 * (I.noff[body] + n1);
      for (int k = 0; k < 3; ++k) lists[g0 + k].push_back(m);
      for (int k = 0; k < 3; ++k) lists[g1 + k].push_back(m);
      for (int k = 0; k < 3; ++k) lists[g0 + 3 + k].push_back(m);
      for (int k = 0; k < 3; ++k) lists[g1 + 3 + k].push_back(m);
    }
  std::vector<int> ptr(I.ndof + 1, 0), lst, kind(I.ndof);
  for (int d = 0; d < I.ndof; ++d) {
    ptr[d + 1] = ptr[d] + (int)lists[d].size();
    lst.insert(lst.end(), lists[d].begin(), lists[d].end());
    kind[d] = d % 6;
  }
  if (upload(I.dof_ptr, ptr.data(), ptr.size(), b->stream)) return 1;
  if (upload(I.dof_list, lst.data(), lst.size(), b->stream)) return 1;
  {
    std::vector<int> lel(lst.size());  // the element of each list entry's marker
    for (size_t j = 0; j < lst.size(); ++j) lel[j] = I.marker_elem[lst[j]];
    if (upload(I.dof_elem, lel.data(), lel.size(), b->stream)) return 1;
  }
  if (upload(I.dof_kind, kind.data(), kind.size(), b->stream)) return 1;
  if (upload(I.dmarker_elem, I.marker_elem.data(), I.marker_elem.size(), b->stream)) return 1;
  HIPCHK(I.Exyz.ensure(3 * (size_t)I.M));
  HIPCHK(I.Evel.ensure(3 * (size_t)I.M));
  HIPCHK(I.Ea.ensure(I.M));
  HIPCHK(I.Eforce.ensure(3 * (size_t)I.M));
  HIPCHK(I.fT.ensure(3 * (size_t)I.M));
  HIPCHK(I.lod.ensure(I.ndof));
  HIPCHK(I.elem_xc.ensure(3 * (size_t)std::max(1, I.nelems)));
  // device stencil and band structures
  std::vector<int> first(std::max(1, I.M));
  for (int body = 0; body < nbodies; ++body)
    for (int m = I.moff[body]; m < I.moff[body + 1]; ++m) first[m] = I.moff[body];
  if (upload(I.first_marker, first.data(), first.size(), b->stream)) return 1;
  const int E = 64 * I.M;
  I.cap = std::max(1, E);
  if (b->N > (1LL << 32)) return fail("block too large for the band keys (at most 2^32 cells)");
  if (64LL * I.M >= (1LL << 31)) return fail("too many markers for the band (64 M must be below 2^31)");
  HIPCHK(I.dEw.ensure(12 * (size_t)std::max(1, I.M)));
  HIPCHK(I.keys.ensure(I.cap));
  HIPCHK(I.tmp.ensure(I.cap));
  HIPCHK(I.blkbox.ensure(6 * (size_t)((I.cap + kClaimThreads - 1) / kClaimThreads)));
  HIPCHK(I.flags.ensure(I.cap));
  HIPCHK(I.scan.ensure((size_t)I.cap + kScanRun));  // k_band_scan reads whole runs of kScanRun
  HIPCHK(I.dentry.ensure(I.cap));
  HIPCHK(I.went.ensure(I.cap));
  HIPCHK(I.dbidx.ensure(I.cap));
  HIPCHK(I.dptr.ensure((size_t)I.cap + 1));
  HIPCHK(I.dcell.ensure(I.cap));
  HIPCHK(I.band_u.ensure(3 * (size_t)I.cap));
  HIPCHK(I.band_F.ensure(3 * (size_t)I.cap));
  HIPCHK(I.dstate.ensure(4));
  HIPCHK(I.tolpart.ensure(std::max(1, I.M)));  // per-marker |dV|_1
  HIPCHK(I.ctl.ensure(kCtlWords));
  HIPCHK(I.cellmap.ensure((size_t)b->N));
  {
    const size_t m1 = std::max(1, I.M);
    HIPCHK(I.wsX.ensure(12 * m1));
    HIPCHK(I.wsY.ensure(3 * m1));
    HIPCHK(I.wsP.ensure(3 * m1));
    HIPCHK(I.Eprev.ensure(3 * m1));
    HIPCHK(I.Enrm.ensure(3 * m1));
    HIPCHK(I.Esnap.ensure(3 * m1));
    HIPCHK(hipMemsetAsync(I.Eprev.p, 0, 3 * m1 * sizeof(double), b->stream));
    HIPCHK(hipMemsetAsync(I.Enrm.p, 0, 3 * m1 * sizeof(double), b->stream));
    // dstate: residual, (unused), the last call's residual at iteration skip+1 (-1: none), (unused)
    const double d0[4] = {1e10, 0.0, -1.0, 0.0};
    if (upload(I.dstate, d0, 4, b->stream)) return 1;
  }
  HIPCHK(hipMemsetAsync(I.ctl.p, 0, kCtlWords * sizeof(int), b->stream));
  HIPCHK(hipMemsetAsync(I.cellmap.p, 0xff, (size_t)b->N * sizeof(int), b->stream));  // -1
  HIPCHK(I.rep->ensure());
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_ibm_set_params(int h, int ntolLBM, double dtolLBM, double Uref, double denIn) {
  Block *b = get_block(h);
  if (!b) return 1;
  b->ibm.ntol = ntolLBM;
  b->ibm.dtol = dtolLBM;
  b->ibm.Uref = Uref;
  b->ibm.denIn = denIn;
  return 0;
}

int fsil_ibm_interaction_force(int h, double dt, const double *Exyz, const double *Evel,
                               const double *Ea, const double *elem_xc, double *Eforce,
                               double *lodFlow, int *iters, double *dmax) {
  Block *b = get_block(h);
  if (!b) return 1;
  Ibm &I = b->ibm;
  if (ibm_check_report(*b)) return 1;  // an earlier call that did not wait
  I.active = false;
  if (I.nbodies == 0 || I.M == 0) {
    if (iters) *iters = 0;
    if (dmax) *dmax = 0.0;
    return 0;
  }
  hipStream_t s = coupling_stream(*b);
  const int M = I.M, E = 64 * M;
  // workgroups of the band-cell kernels (kCellsPerPass cells per pass): the
  // last call's band and a quarter more, at most one per kCellsPerPass entries
  const int cell_cap = (E + kCellsPerPass - 1) / kCellsPerPass;
  const int cell_blocks =
      I.last_nband > 0
          ? std::min(cell_cap, (I.last_nband + I.last_nband / 4 + kCellsPerPass - 1) / kCellsPerPass + 1)
          : cell_cap;
  if (fem_order_after(*b)) return 1;  // k_ibm_finish rewrites the loads the beam sub-steps read
  if (!Exyz) {
    // markers computed on the device by fsil_fem_update_markers
    if (!b->fem.markers_on_device) return fail("no marker arrays given and no device markers computed");
  } else {
    if (!Evel || !Ea || !elem_xc) return fail("null marker arrays");
    if (upload(I.Exyz, Exyz, 3 * (size_t)M, s)) return 1;
    if (upload(I.Evel, Evel, 3 * (size_t)M, s)) return 1;
    if (upload(I.Ea, Ea, (size_t)M, s)) return 1;
    if (upload(I.elem_xc, elem_xc, 3 * (size_t)I.nelems, s)) return 1;
  }
  // UpdateElmtInterp_ (Solidbody.f90:760-867) and the band, on the device
  const int done0 = (I.ntol > 0 && 1e10 > I.dtol) ? 0 : 1;
  Mlp *net = nullptr;
  if (I.warm_mlp) {
    auto it = g_mlps.find(I.warm_mlp);
    if (it == g_mlps.end()) return fail("warm-start network %d no longer exists", I.warm_mlp);
    net = it->second.get();
    // the network's parameters are written on its own stream (training)
    HIPCHK(hipStreamSynchronize(net->s));
  }
  // the guess needs the previous call's force (Eprev) and the skip within the loop's cap
  const int warm = (net && !done0 && I.have_prev && I.skip < I.ntol) ? 1 : 0;
  hipLaunchKernelGGL(k_ibm_reset, dim3(std::min(1024, (E + 255) / 256)), dim3(256), 0, s, I.ctl.p, I.dcell.p,
                     I.cellmap.p, I.dstate.p, done0, warm, I.skip, E, I.scan.p, I.dptr.p, I.Eforce.p, 3 * M);
  HIPCHK(hipGetLastError());
  StencilArgs sa{};
  sa.M = M;
  sa.Exyz = I.Exyz.p;
  sa.first_marker = I.first_marker.p;
  sa.org[0] = b->xmin; sa.org[1] = b->ymin; sa.org[2] = b->zmin;
  sa.dh = b->dh;
  sa.invdh = 1.0 / b->dh;
  sa.nd[0] = b->xd; sa.nd[1] = b->yd; sa.nd[2] = b->zd;
  for (int k = 0; k < 6; ++k) sa.bc[k] = b->bc[k];
  sa.Ew = I.dEw.p;
  sa.keys = I.keys.p;
  sa.ctl = I.ctl.p;
  // band: cells bid for by the stencils and claimed, per-cell counts (I.scan) and slots
  // (I.flags), CSR pointers by an inclusive scan into dptr[1..], entries
  // placed in arrival order (I.tmp), then at their ranks by entry number
  sa.cellmap = I.cellmap.p;  // (the stencil kernel makes the bids)
  hipLaunchKernelGGL(k_ibm_stencil, dim3((M + 3) / 4), dim3(256), 0, s, sa);
  const int nclaim = (E + kClaimThreads - 1) / kClaimThreads;
  hipLaunchKernelGGL(k_ibm_claim, dim3(nclaim), dim3(kClaimThreads), 0, s, E, I.keys.p, I.cellmap.p, I.dcell.p,
                     b->yd, b->zd, I.blkbox.p, I.ctl.p);
  HIPCHK(hipGetLastError());
  // calculate_interaction_force (Solidbody.f90:869-918): band velocity of the stored state with the count
  const double hv0 = 0.5 * b->VF[0] * b->dh, hv1 = 0.5 * b->VF[1] * b->dh, hv2 = 0.5 * b->VF[2] * b->dh;
  BandMacroArgs ma{};
  ma.f = b->precision == 0 ? b->fA.p : nullptr;
  ma.g = b->precision == 0 ? nullptr : b->gA.p;
  ma.N = b->N; ma.cell = I.dcell.p; ma.cap = I.cap; ma.hv0 = hv0; ma.hv1 = hv1; ma.hv2 = hv2; ma.band_u = I.band_u.p;
  hipLaunchKernelGGL(k_ibm_count, dim3((E + 255) / 256), dim3(256), 0, s, E, I.keys.p, I.cellmap.p, I.dbidx.p,
                     I.scan.p, I.flags.p, I.ctl.p, ma);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_band_scan, dim3(1), dim3(1024), 0, s, I.ctl.p, I.scan.p, I.dptr.p, I.blkbox.p, nclaim);
  hipLaunchKernelGGL(k_ibm_fill, dim3((E + 255) / 256), dim3(256), 0, s, E, I.dbidx.p, I.flags.p, I.dptr.p,
                     I.tmp.p);
  hipLaunchKernelGGL(k_ibm_rank, dim3((E + 255) / 256), dim3(256), 0, s, E, I.ctl.p, I.dbidx.p, I.dptr.p, I.tmp.p,
                     I.dentry.p, I.dEw.p, I.went.p);
  HIPCHK(hipGetLastError());
  const double invh3p = 0.5 * dt * ((1.0 / b->dh) * (1.0 / b->dh) * (1.0 / b->dh)) / I.denIn;
  double dsum = 0.0;
  for (int body = 0; body < I.nbodies; ++body) dsum = dsum + (double)I.nmark[body];
  if (warm || I.record) {
    hipLaunchKernelGGL(k_ws_features, dim3((M + 3) / 4), dim3(256), 0, s, M, I.dbidx.p, I.dEw.p, I.band_u.p, I.cap,
                       I.Evel.p, I.Ea.p, I.Eprev.p, I.Enrm.p, I.Uref, I.wsX.p);
    HIPCHK(hipGetLastError());
  }
  if (warm) {
    // the learned guess of the force after `skip` rounds, applied to the band
    // as one correction of the uncorrected velocity
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((M + 63) / 64), dim3(256), 0, s, I.wsX.p, M, net->th.p, I.wsP.p);
    hipLaunchKernelGGL(k_ws_apply, dim3((3 * M + 255) / 256), dim3(256), 0, s, M, I.wsP.p, I.Ea.p, I.Eprev.p, I.Uref,
                       invh3p, I.Eforce.p, I.fT.p);
    hipLaunchKernelGGL(k_correct, dim3(cell_blocks), dim3(256), 0, s, I.ctl.p, -1, MdfTest{}, I.cap, I.dptr.p,
                       I.dentry.p, I.went.p, I.fT.p, I.band_u.p, cell_blocks, I.pass_cap);
    HIPCHK(hipGetLastError());
  }
  // MDF loop (Solidbody.f90:893-906): the rounds are queued (with the warm
  // start one more than the cap, in case it falls back and restarts); a round
  // is k_interp and k_correct, whose last workgroup makes the stop test, and
  // the kernels of the rounds after the stop return at once.  A plain call
  // that records training data keeps the force after `skip` rounds (Esnap).
  const int rounds = I.ntol + warm;
  const int snap = (!warm && I.record && I.skip > 0) ? I.skip - 1 : -1;
  MdfTest mt{};
  mt.tl = I.tolpart.p; mt.dstate = I.dstate.p; mt.M = M; mt.ntol = I.ntol;
  mt.dsum = dsum; mt.Uref = I.Uref; mt.dtol = I.dtol; mt.skip = I.skip;
  for (int it = 0; it < rounds; ++it) {
    hipLaunchKernelGGL(k_interp, dim3((M + 3) / 4), dim3(256), 0, s, M, it, I.dbidx.p, I.dEw.p, I.band_u.p, I.cap,
                       I.ctl.p, I.Evel.p, I.Ea.p, invh3p, I.Eforce.p, I.fT.p, I.tolpart.p, snap, I.Esnap.p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_correct, dim3(cell_blocks + 1), dim3(256), 0, s, I.ctl.p, it, mt, I.cap, I.dptr.p,
                       I.dentry.p, I.went.p, I.fT.p, I.band_u.p, cell_blocks, I.pass_cap);
    HIPCHK(hipGetLastError());
    if (warm && it == 0) {
      // restart from the zero guess when the first round asked for it
      hipLaunchKernelGGL(k_ws_fallback, dim3((3 * M + 255) / 256), dim3(256), 0, s, M, I.ctl.p, I.Eforce.p);
      if (b->precision == 0)
        hipLaunchKernelGGL(k_band_macro, dim3((E + 255) / 256), dim3(256), 0, s, b->fA.p, b->N, I.dcell.p,
                           I.ctl.p, I.cap, hv0, hv1, hv2, I.band_u.p, 1);
      else
        hipLaunchKernelGGL(k_band_macro32, dim3((E + 255) / 256), dim3(256), 0, s, b->gA.p, b->N, I.dcell.p,
                           I.ctl.p, I.cap, hv0, hv1, hv2, I.band_u.p, 1);
      hipLaunchKernelGGL(k_ws_fallback_end, dim3(1), dim3(64), 0, s, I.ctl.p);
      HIPCHK(hipGetLastError());
    }
  }
  if (I.warm_mlp || I.record) {
    // training targets of a plain call; the previous force of the next call
    hipLaunchKernelGGL(k_ws_targets, dim3((3 * M + 255) / 256), dim3(256), 0, s, M, I.ctl.p, I.skip, I.Eforce.p,
                       I.Esnap.p, I.Ea.p, I.Uref, (I.record && !warm) ? 1 : 0, I.wsY.p, I.Eprev.p);
    HIPCHK(hipGetLastError());
  }
  I.have_prev = true;
  // spreading (Solidbody.f90:920-979), beam loads and the report in one launch
  IbmFinishArgs fa{};
  fa.ctl = I.ctl.p; fa.dstate = I.dstate.p; fa.rep = I.rep->d;
  fa.cap = I.cap; fa.ptr = I.dptr.p; fa.entry = I.dentry.p; fa.went = I.went.p; fa.Eforce = I.Eforce.p;
  fa.invh3 = (1.0 / b->dh) * (1.0 / b->dh) * (1.0 / b->dh);
  fa.VF1 = b->VF[0]; fa.VF2 = b->VF[1]; fa.VF3 = b->VF[2];
  fa.band_F = I.band_F.p;
  fa.ndof = I.ndof; fa.dof_blocks = (I.ndof + 3) / 4;
  fa.dof_ptr = I.dof_ptr.p; fa.dof_list = I.dof_list.p; fa.dof_elem = I.dof_elem.p; fa.dof_kind = I.dof_kind.p;
  fa.marker_elem = I.dmarker_elem.p; fa.elem_xc = I.elem_xc.p; fa.Exyz = I.Exyz.p;
  fa.lod = I.lod.p;
  fa.band_blocks = cell_blocks;
  fa.pass_cap = I.pass_cap;
  hipLaunchKernelGGL(k_ibm_finish, dim3(fa.dof_blocks + cell_blocks), dim3(256), 0, s, fa);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(I.rep->ev, s));
  I.rep->pending = true;
  if (!(Eforce || lodFlow || iters || dmax)) {
    // nothing asked back: no host wait; the report is checked by the next
    // synchronising coupling call (an error then surfaces there)
    I.active = true;
    return 0;
  }
  if (Eforce) HIPCHK(hipMemcpyAsync(Eforce, I.Eforce.p, 3 * (size_t)M * sizeof(double), hipMemcpyDeviceToHost, s));
  if (lodFlow && I.ndof > 0) HIPCHK(hipMemcpyAsync(lodFlow, I.lod.p, (size_t)I.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (fem_collect(*b)) return 1;  // results of the previous beam sub-steps
  if (ibm_check_report(*b)) return 1;
  I.active = true;
  if (iters) *iters = I.last_iters;
  if (dmax) *dmax = I.last_dmax;
  return 0;
}

int fsil_ibm_forces(int h, double *Eforce, double *lodFlow) {
  Block *b = get_block(h);
  if (!b) return 1;
  Ibm &I = b->ibm;
  if (I.M == 0) return 0;
  if (Eforce) HIPCHK(hipMemcpyAsync(Eforce, I.Eforce.p, 3 * (size_t)I.M * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  if (lodFlow && I.ndof > 0)
    HIPCHK(hipMemcpyAsync(lodFlow, I.lod.p, (size_t)I.ndof * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return ibm_check_report(*b);
}

int fsil_ibm_report(int h, int *iters, double *dmax) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (ibm_check_report(*b)) return 1;
  if (iters) *iters = b->ibm.last_iters;
  if (dmax) *dmax = b->ibm.last_dmax;
  return 0;
}

int fsil_ibm_set_warm_start(int h, int mlp, int record, int skip) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (mlp && !get_mlp(mlp)) return 1;
  Ibm &I = b->ibm;
  if (skip < 0 || (I.ntol > 0 && skip >= I.ntol)) return fail("warm-start skip %d outside [0, ntolLBM)", skip);
  if (mlp && !I.warm_mlp && !I.record && I.M > 0) {
    // the previous force of the next call: the last call's final marker force
    HIPCHK(hipMemcpyAsync(I.Eprev.p, I.Eforce.p, 3 * (size_t)I.M * sizeof(double), hipMemcpyDeviceToDevice,
                          b->stream));
  }
  I.warm_mlp = mlp;
  I.record = record ? 1 : 0;
  I.skip = skip;
  return 0;
}

int fsil_ibm_rounds(int h, int *rounds) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (ibm_check_report(*b)) return 1;
  if (rounds) *rounds = b->ibm.last_rounds;
  return 0;
}

int fsil_ibm_warm_start_data(int h, float *X, float *Y, int *info) {
  Block *b = get_block(h);
  if (!b) return 1;
  Ibm &I = b->ibm;
  if (!(I.record || I.warm_mlp)) return fail("warm-start features are computed only with a network or recording on");
  if (X) HIPCHK(hipMemcpyAsync(X, I.wsX.p, 12 * (size_t)I.M * sizeof(float), hipMemcpyDeviceToHost, b->stream));
  if (Y) HIPCHK(hipMemcpyAsync(Y, I.wsY.p, 3 * (size_t)I.M * sizeof(float), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  if (ibm_check_report(*b)) return 1;
  if (info) {
    info[0] = I.last_warm;
    info[1] = I.last_fallback;
  }
  return 0;
}

}  // extern "C"

namespace {
// The host's check of a coupling call's report (waits for it if needed):
// the stencil, NaN and band errors of the reference, then the iteration
// count and residual of the call.
int ibm_check_report(Block &b) {
  Ibm &I = b.ibm;
  if (!I.rep || !I.rep->pending) return 0;
  I.rep->pending = false;
  HIPCHK(step_wait(I.rep->ev));
  const IbmReport &R = *I.rep->h;
  const int err = R.ctl[kErr];
  if (err != 0x7fffffff) {
    I.active = false;
    if (err == 3) return fail("Nan found in PenaltyForce");
    if (err == 4) return fail("immersed-boundary band: a stencil cell without a band number");
    const int m = err / 4 - 1, why = err % 4;
    int body = 0;
    while (body + 1 < I.nbodies && m >= I.moff[body + 1]) ++body;
    return fail("immersed-boundary stencil of marker %d (body %d): index out of %s bound", m - I.moff[body] + 1,
                body + 1, why == 1 ? "min" : "max");
  }
  I.have_box = R.ctl[kNband] > 0;
  for (int k = 0; k < 6; ++k) I.box[k] = R.ctl[kBoxLo + k];
  I.last_warm = R.ctl[kGuess] && !R.ctl[kFallback];  // the guess was kept
  I.last_fallback = R.ctl[kFallback] != 0;
  // forcing rounds run: interpolation rounds plus the guess's correction
  I.last_rounds = R.ctl[kRan] + 1 + R.ctl[kGuess];
  I.last_nband = R.ctl[kNband];
  I.last_iters = R.ctl[kIters];
  I.last_dmax = R.dmax;
  return 0;
}

// Wait (host side) for the last beam sub-steps and check their results:
// "CG solver breakdown" of any sub-step and body is an error.
int fem_collect(Block &b) {
  Fem &F = b.fem;
  if (!F.fs || !F.fs->pending) return 0;
  HIPCHK(step_wait(F.fs->ev_done));
  F.fs->pending = false;
  const int nb = F.nbodies;
  const int nsub = nb ? F.fs->n_staged / nb : 0;
  for (int i = 0; i < F.fs->n_staged; ++i)
    if (F.fs->h_stage[i].status)
      return fail("ERROR: CG solver breakdown (body %d, sub-step %d)", i % nb + 1, i / nb + 1);
  if (nsub > 0)
    for (int bi = 0; bi < nb; ++bi) {
      const FemBodyDesc &r = F.fs->h_stage[(nsub - 1) * nb + bi];
      F.desc[bi].newton_iters = r.newton_iters;
      F.desc[bi].cg_iters = r.cg_iters;
      F.desc[bi].status = r.status;
      F.desc[bi].dnorm = r.dnorm;
    }
  return 0;
}

// Kernel arguments of k_solid_info for body `body`, with `buf` (3 nEL + 78
// doubles of device memory) for the per-element energies and the rows.  The
// node vectors and loads are those of the last beam sub-steps and coupling
// call (ordered before on the block's stream by the caller's waits).
int fill_solid_info_args(Block &b, int body, const double *refs, double *buf, SolidInfoArgs &a) {
  Fem &F = b.fem;
  Ibm &I = b.ibm;
  const FemBodyDesc &d = F.desc[body];
  if (I.ndof != 6 * F.nND) return fail("immersed-boundary bodies do not match the beam bodies");
  a.el = F.el.p + d.eoff;
  a.pos = F.pos.p + d.goff;
  a.vel = F.vel.p + d.goff;
  a.acc = F.acc.p + d.goff;
  a.lod = I.lod.p + d.goff;
  a.nND = d.nND;
  a.nEL = d.nEL;
  for (int k = 0; k < 3; ++k) a.XYZo[k] = F.bh[body].XYZo[k];
  for (int k = 0; k < 6; ++k) a.refs[k] = refs[k];
  a.parts = buf;
  a.rows = buf + 3 * (size_t)d.nEL;
  return 0;
}

int collide_any(Block &b, int region, const int *hole) {
  return b.precision == 0 ? collide_stream(b, region, hole) : collide_stream32(b, region, hole);
}

// Rows (x0, x1, y0, y1) of the predicted band box of a step's coupling call
// for a split step: the last checked call's band box widened by
// split_margin cells (3: markers move well under a cell per step; a band
// that leaves the prediction is still collided by the hull launch), y
// aligned to the collide workgroup's rows so that a workgroup is skipped or
// covered whole.  False when the step is not split: no earlier band box, no
// beam stream, the LES models (WALE reads the relaxation time that a cell's
// collision rewrites), or FSIL_SPLIT=0.
bool split_hole(const Block &b, int *hole) {
  const Ibm &I = b.ibm;
  if (!b.split_ok || !I.have_box || I.M == 0 || is_les(b.model) || !b.fem.fs || !b.fem.fs->s) return false;
  const int by = (int)(b.precision == 1 ? cell_block32(b) : cell_block(b)).y;
  const int m = b.split_margin;
  const int x0 = std::max(0, I.box[0] - m), x1 = std::min(b.xd - 1, I.box[3] + m);
  int y0 = std::max(0, I.box[1] - m), y1 = std::min(b.yd - 1, I.box[4] + m);
  if (x1 < x0 || y1 < y0) return false;
  y0 = y0 / by * by;
  y1 = std::min(b.yd - 1, (y1 / by + 1) * by - 1);
  hole[0] = x0; hole[1] = x1; hole[2] = y0; hole[3] = y1;
  return true;
}

// order work on the block's stream after the last beam sub-steps (device side)
int fem_order_after(Block &b) {
  Fem &F = b.fem;
  if (b.split) return 0;  // the coupling runs on the beam stream itself
  if (F.fs && F.fs->pending) HIPCHK(hipStreamWaitEvent(b.stream, F.fs->ev_done, 0));
  return 0;
}
}  // namespace

extern "C" {

int fsil_fem_define_bodies(int h, int nbodies, const int *nND, const int *nEL, const double *xyz8,
                           const int *elems, const int *cons, const int *iBodyModel, const int *isMotionGiven,
                           const double *mat, const double *kin, double Lref, double Uref, double denIn,
                           double alpha) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (nbodies <= 0) return fail("fsil_fem_define_bodies needs at least one body");
  Fem &F = b->fem;
  F = Fem();
  F.nbodies = nbodies;
  std::vector<BeamElem> all_el;
  std::vector<double> all_pos, all_vel;
  std::vector<int> nmark(nbodies), me, en, fixed;
  std::vector<std::vector<std::pair<int, int>>> dofl;
  std::vector<std::vector<std::pair<int, int>>> nodel;
  std::vector<int> body_of_elem, first_marker, goff(nbodies);
  std::vector<double> beta(nbodies);
  size_t xo = 0, eo = 0, co = 0;
  int mcount = 0;
  F.desc.assign(nbodies, FemBodyDesc{});
  F.bh.assign(nbodies, FemBodyHost{});
  {
    const char *e = getenv("FSIL_FEM_HOT");
    F.hot_ok = !(e && !strcmp(e, "0"));
  }
  for (int bi = 0; bi < nbodies; ++bi) {
    if (nND[bi] < 2 || nEL[bi] < 1) return fail("body %d: needs nodes and elements", bi + 1);
    if (nND[bi] > 510)  // the solver keeps a body's CG products in one LDS staging area
      return fail("body %d: the device beam solver takes bodies of at most 510 nodes (%d given)", bi + 1, nND[bi]);
    for (int e = 0; e < nEL[bi]; ++e)
      for (int k = 0; k < 2; ++k) {
        const int nd = elems[3 * (eo + e) + k];
        if (nd < 1 || nd > nND[bi]) return fail("body %d element %d: node %d out of range", bi + 1, e + 1, nd);
      }
    std::vector<BeamElem> els;
    std::vector<double> pos, vel;
    fem_init_body(xyz8 + 8 * xo, nND[bi], nEL[bi], elems + 3 * eo, cons + 6 * xo, iBodyModel[bi],
                  isMotionGiven + 6 * bi, mat + 4 * bi, kin + 22 * bi, Lref, Uref, denIn, F.bh[bi], els, pos, vel);
    for (int e = 0; e < nEL[bi]; ++e) {
      if (const int k = el_stiff_scalars(els[e]))
        return fail("body %d element %d: local stiffness entry %d off its pattern", bi + 1, e + 1, k);
      if (el_mass_blocks_repeat(els[e])) F.hot_ok = false;
    }
    FemBodyDesc &d = F.desc[bi];
    d.nND = nND[bi]; d.nEL = nEL[bi]; d.eoff = F.nEL; d.noff = F.nND; d.goff = 6 * F.nND;
    d.iBodyModel = iBodyModel[bi];
    goff[bi] = d.goff;
    beta[bi] = -(alpha * 2.0 * denIn);
    const int g0 = (int)fixed.size();
    fixed.resize(g0 + 6 * (size_t)nND[bi], 0);
    dofl.resize(g0 + 6 * (size_t)nND[bi]);
    nodel.resize(F.nND + (size_t)nND[bi]);
    int mb = 0;
    for (int e = 0; e < nEL[bi]; ++e) {
      const BeamElem &E = els[e];
      const int ge = F.nEL + e;
      for (int i = 0; i < NED; ++i) {
        const int gl = i < 6 ? 6 * E.node0 + i : 6 * E.node1 + (i - 6);
        dofl[g0 + gl].push_back({ge, i});
        if (E.bc[i] > 0) fixed[g0 + gl] = 1;
      }
      nodel[F.nND + E.node0].push_back({ge, 0});
      nodel[F.nND + E.node1].push_back({ge, 1});
      body_of_elem.push_back(bi);
      first_marker.push_back(mcount + mb);
      for (int s2 = 0; s2 < E.Nspan; ++s2) me.push_back(e);
      en.push_back(E.node0);
      en.push_back(E.node1);
      mb += E.Nspan;
    }
    nmark[bi] = mb;
    mcount += mb;
    all_el.insert(all_el.end(), els.begin(), els.end());
    all_pos.insert(all_pos.end(), pos.begin(), pos.end());
    all_vel.insert(all_vel.end(), vel.begin(), vel.end());
    F.nEL += nEL[bi];
    F.nND += nND[bi];
    xo += nND[bi];
    eo += nEL[bi];
    co += 6 * (size_t)nND[bi];
  }
  F.gEQ = 6 * F.nND;
  // the immersed-boundary side of the same bodies
  std::vector<int> nn(nND, nND + nbodies), ne(nEL, nEL + nbodies);
  if (fsil_ibm_define_bodies(h, nbodies, nmark.data(), nn.data(), ne.data(), me.data(), en.data())) return 1;
  std::vector<int> dptr(F.gEQ + 1, 0), de, di, nptr(F.nND + 1, 0), ne2, nh;
  for (int g = 0; g < F.gEQ; ++g) {
    if (dofl[g].empty()) return fail("degree of freedom %d belongs to no element", g);
    dptr[g + 1] = dptr[g] + (int)dofl[g].size();
    for (auto &pr : dofl[g]) { de.push_back(pr.first); di.push_back(pr.second); }
  }
  for (int n = 0; n < F.nND; ++n) {
    nptr[n + 1] = nptr[n] + (int)nodel[n].size();
    for (auto &pr : nodel[n]) { ne2.push_back(pr.first); nh.push_back(pr.second); }
  }
  hipStream_t s = b->stream;
  std::vector<double> zeros(F.gEQ, 0.0);
  if (upload(F.el, all_el.data(), all_el.size(), s)) return 1;
  if (upload(F.pos, all_pos.data(), all_pos.size(), s)) return 1;
  if (upload(F.vel, all_vel.data(), all_vel.size(), s)) return 1;
  if (upload(F.dsp, zeros.data(), zeros.size(), s)) return 1;
  if (upload(F.acc, zeros.data(), zeros.size(), s)) return 1;
  if (upload(F.fixed, fixed.data(), fixed.size(), s)) return 1;
  if (upload(F.dof_ptr, dptr.data(), dptr.size(), s)) return 1;
  if (upload(F.dof_e, de.data(), de.size(), s)) return 1;
  if (upload(F.dof_i, di.data(), di.size(), s)) return 1;
  if (upload(F.node_ptr, nptr.data(), nptr.size(), s)) return 1;
  if (upload(F.node_e, ne2.data(), ne2.size(), s)) return 1;
  if (upload(F.node_h, nh.data(), nh.size(), s)) return 1;
  if (upload(F.body_of_elem, body_of_elem.data(), body_of_elem.size(), s)) return 1;
  if (upload(F.elem_first_marker, first_marker.data(), first_marker.size(), s)) return 1;
  if (upload(F.body_goff, goff.data(), goff.size(), s)) return 1;
  if (upload(F.beta, beta.data(), beta.size(), s)) return 1;
  for (DevBuf<double> *v : {&F.dspO, &F.velO, &F.accO, &F.lodInte, &F.lodEffe, &F.lodExte, &F.vBC, &F.x, &F.r, &F.p,
                            &F.Ap, &F.z, &F.xF})
    HIPCHK(v->ensure(F.gEQ));
  HIPCHK(F.lbuf.ensure((size_t)F.nEL * NED));
  HIPCHK(F.blk.ensure((size_t)F.nND * 36));
  F.fs = std::make_unique<FemStream>();
  HIPCHK(hipStreamCreateWithFlags(&F.fs->s, hipStreamNonBlocking));
  {
    int dev = 0, lim = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    F.lds_limit = (size_t)lim;
  }
  HIPCHK(hipEventCreateWithFlags(&F.fs->ev_in, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&F.fs->ev_done, hipEventDisableTiming));
  // lodFlow starts at zero (Beam_InitLoad)
  if (upload(b->ibm.lod, zeros.data(), zeros.size(), s)) return 1;
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int fsil_fem_set_solver(int h, double dampK, double dampM, double GeoGamma, double NewmarkGamma,
                        double NewmarkBeta, double dtolFEM, int ntolFEM) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  F.dampK = dampK; F.dampM = dampM; F.GeoGamma = GeoGamma; F.NewmarkGamma = NewmarkGamma;
  F.NewmarkBeta = NewmarkBeta; F.dtolFEM = dtolFEM; F.ntolFEM = ntolFEM;
  return 0;
}

int fsil_fem_update_markers(int h) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  Ibm &I = b->ibm;
  if (F.nbodies == 0) return fail("no FEM bodies defined");
  if (fem_order_after(*b)) return 1;
  FemMarkerArgs a{};
  a.el = F.el.p; a.pos = F.pos.p; a.vel = F.vel.p;
  a.body_of_elem = F.body_of_elem.p; a.elem_first_marker = F.elem_first_marker.p; a.body_goff = F.body_goff.p;
  a.beta = F.beta.p; a.nEL_total = F.nEL;
  a.Exyz = I.Exyz.p; a.Evel = I.Evel.p; a.Ea = I.Ea.p; a.elem_xc = I.elem_xc.p; a.Enrm = I.Enrm.p;
  hipLaunchKernelGGL(k_fem_markers, dim3((F.nEL + 3) / 4), dim3(256), 0, coupling_stream(*b), a);
  HIPCHK(hipGetLastError());
  F.markers_on_device = true;
  return 0;
}

int fsil_fem_substeps(int h, double time, int numsubstep, double deltat) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  if (F.nbodies == 0) return fail("no FEM bodies defined");
  if (numsubstep < 1) return fail("numsubstep must be >= 1");
  if (!F.fs || !F.fs->s) return fail("FEM bodies not fully defined (fsil_fem_define_bodies failed)");
  FemStream &S = *F.fs;
  if (fem_collect(*b)) return 1;  // the staging buffer is free once the previous call is done
  const double subdt = deltat / (double)numsubstep;
  FemArrays A = fem_arrays(*b);
  int max_nd = 1;
  for (const FemBodyDesc &d : F.desc) max_nd = std::max(max_nd, d.nND);
  int max_ne = 1;
  for (const FemBodyDesc &d : F.desc) max_ne = std::max(max_ne, d.nEL);
  // LDS: the work area (dot-product staging, and the CG's assembly lists when
  // they fit), the CG vectors and flags, the element matrices and the element
  // fields, each as far as the device's workgroup LDS holds them (beyond
  // 64 KiB after raising the kernel's dynamic-LDS limit)
  const size_t lds_lim = F.lds_limit > 1024 ? F.lds_limit - 1024 : 64 * 1024;
  auto allow = [&](int hot, size_t bytes) {
    if (bytes <= F.lds_set[hot]) return true;
    const void *k = hot ? (const void *)k_fem_substep<true> : (const void *)k_fem_substep<false>;
    if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess) {
      F.lds_set[hot] = bytes;
      return true;
    }
    (void)hipGetLastError();
    return false;
  };
  size_t cg_need = 0;
  for (const FemBodyDesc &d : F.desc) cg_need = std::max(cg_need, fem_cg_work(d.nND, d.nEL));
  // 2 nG of dot-product staging, >= one maximum per wave, >= the six end-rotation values per element of the
  // element-stress phase
  const int stage = std::max(std::max(12 * max_nd, 6 * max_ne), 16);
  // (the areas' offsets are even, from the kernel's 16-byte aligned base)
  int cg_lds = 1, work = ((int)std::max(cg_need, (size_t)stage) + 1) & ~1;
  if (fem_lds_bytes(work, 1, max_nd, max_ne) > lds_lim) {
    cg_lds = 0;
    work = (stage + 1) & ~1;
  }
  int vec_lds = 0, coef_off = 0, hot_off = 0;
  size_t lds_bytes = fem_lds_bytes(work, 0, max_nd, max_ne);
  {
    const size_t b1 = fem_lds_bytes(work, 1, max_nd, max_ne);
    if (b1 <= lds_lim && allow(0, b1)) {
      vec_lds = 1;
      lds_bytes = b1;
      const size_t b2 = b1 + (size_t)max_ne * kCoefStride * sizeof(double);
      if (b2 <= lds_lim && allow(0, b2)) {
        coef_off = (int)(b1 / sizeof(double)) - 1;  // past the slack double
        lds_bytes = b2;
        const size_t b3 = b2 + (size_t)max_ne * sizeof(BeamHot);
        if (F.hot_ok && F.dampK <= 0.0 && b3 <= lds_lim && allow(1, b3)) {
          hot_off = (int)(b2 / sizeof(double)) - 1;
          lds_bytes = b3;
        }
      }
    }
  }
  const size_t n = (size_t)numsubstep * F.nbodies;
  if (n > S.cap) {
    if (S.h_stage) HIPCHK(hipHostFree(S.h_stage));
    S.h_stage = nullptr;
    S.d_stage = nullptr;
    S.cap = 0;
    HIPCHK(hipHostMalloc((void **)&S.h_stage, n * sizeof(FemBodyDesc), hipHostMallocDefault));
    HIPCHK(hipHostGetDevicePointer((void **)&S.d_stage, S.h_stage, 0));
    S.cap = n;
  }
  for (int isub = 1; isub <= numsubstep; ++isub)
    for (int bi = 0; bi < F.nbodies; ++bi) {
      FemBodyDesc &d = S.h_stage[(size_t)(isub - 1) * F.nbodies + bi];
      d = F.desc[bi];
      fem_kinematics(F.bh[bi], time, isub, deltat, subdt, d);
      double *c = d.coeffs;
      c[2] = 1.0 / (F.NewmarkBeta * subdt);
      c[1] = F.NewmarkGamma * c[2];
      c[0] = c[2] / subdt;
      c[3] = 0.5 / F.NewmarkBeta - 1.0;
      c[4] = F.NewmarkGamma / F.NewmarkBeta - 1.0;
      c[5] = subdt * (0.5 * F.NewmarkGamma / F.NewmarkBeta - 1.0);
      c[6] = subdt * (1.0 - F.NewmarkGamma);
      c[7] = subdt * F.NewmarkGamma;
      d.ntolFEM = F.ntolFEM; d.dtolFEM = F.dtolFEM; d.dampK = F.dampK; d.dampM = F.dampM; d.GeoGamma = F.GeoGamma;
      // FP32-deviation blocks are held to a tolerance: their CG sums its dot products in a tree
      d.tree_sums = b->precision == 1 ? 1 : 0;
      d.newton_iters = 0; d.cg_iters = 0; d.status = 0; d.dnorm = 0.0;
    }
  // the kernel reads the descriptors from the pinned staging area and writes
  // its results there (the previous sub-steps are done with it: fem_collect);
  // it waits for the coupling call that produced lodFlow (block's stream; in
  // a split step the call ran on this stream)
  if (!b->split) {
    HIPCHK(hipEventRecord(S.ev_in, b->stream));
    HIPCHK(hipStreamWaitEvent(S.s, S.ev_in, 0));
  }
  if (hot_off)
    hipLaunchKernelGGL(k_fem_substep<true>, dim3(F.nbodies), dim3(512), lds_bytes, S.s, S.d_stage, F.nbodies,
                       numsubstep, A, work, vec_lds, cg_lds, coef_off, hot_off);
  else
    hipLaunchKernelGGL(k_fem_substep<false>, dim3(F.nbodies), dim3(512), lds_bytes, S.s, S.d_stage, F.nbodies,
                       numsubstep, A, work, vec_lds, cg_lds, coef_off, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(S.ev_done, S.s));
  S.n_staged = (int)n;
  S.pending = true;
  return 0;
}

int fsil_fsi_step(int h, double time, int numsubstep, int *iters, double *dmax) {
  Block *b = get_block(h);
  if (!b) return 1;
  if (b->fem.nbodies <= 0) return fail("fsil_fsi_step needs beam bodies on block %d", h);
  // main.f90:95-97: block time and volume force
  b->blktime = time;
  update_vf(*b);
  int hole[4];
  if (split_hole(*b, hole)) {
    // The collision of the cells outside the predicted band box does not
    // depend on the coupling: it runs on the block's stream while the
    // markers, the coupling call and the beam sub-steps run on the beam
    // stream; the rows of the band box follow the call's report event, then
    // the boundary faces.
    HIPCHK(hipEventRecord(b->ev_fluid, b->stream));
    if (collide_any(*b, kRegionBulk, hole)) return 1;
    HIPCHK(hipStreamWaitEvent(b->fem.fs->s, b->ev_fluid, 0));
    b->split = true;
    int rc = fsil_fem_update_markers(h);
    if (!rc) rc = fsil_ibm_interaction_force(h, b->dh, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                            nullptr);
    if (!rc) rc = fsil_fem_substeps(h, time, numsubstep, b->dh);
    b->split = false;
    if (rc) return 1;
    HIPCHK(hipStreamWaitEvent(b->stream, b->ibm.rep->ev, 0));
    if (collide_any(*b, kRegionHull, hole)) return 1;
    if (launch_faces(*b)) return 1;
    b->ibm.active = false;
    return fsil_ibm_report(h, iters, dmax);
  }
  // IBM_FEM (LBMBlockComm.f90:320-338): markers, the coupling call queued
  // without a host wait, the beam sub-steps queued behind it, then the wait
  // for the call's report, so the beam workgroup is resident before the
  // fluid update's kernels fill the device
  if (fsil_fem_update_markers(h)) return 1;
  if (fsil_ibm_interaction_force(h, b->dh, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr))
    return 1;
  if (fsil_fem_substeps(h, time, numsubstep, b->dh)) return 1;
  if (fsil_ibm_report(h, iters, dmax)) return 1;
  // collision, streaming and boundary faces (LBMBlockComm.f90:285-304)
  return fsil_collide_stream_bc(h);
}

int fsil_fem_state(int h, int body, double *pos, double *vel, double *acc, double *dsp) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  if (body < 0 || body >= F.nbodies) return fail("body %d out of range", body);
  if (fem_collect(*b)) return 1;
  const FemBodyDesc &d = F.desc[body];
  const size_t n = 6 * (size_t)d.nND;
  hipStream_t s = b->stream;
  if (pos) HIPCHK(hipMemcpyAsync(pos, F.pos.p + d.goff, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (vel) HIPCHK(hipMemcpyAsync(vel, F.vel.p + d.goff, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (acc) HIPCHK(hipMemcpyAsync(acc, F.acc.p + d.goff, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (dsp) HIPCHK(hipMemcpyAsync(dsp, F.dsp.p + d.goff, n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int fsil_solid_info(int h, int body, const double *refs, double *rows) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  if (body < 0 || body >= F.nbodies) return fail("body %d out of range", body);
  if (!refs || !rows) return fail("null refs or rows");
  if (fem_collect(*b)) return 1;
  if (ibm_check_report(*b)) return 1;
  const FemBodyDesc &d = F.desc[body];
  HIPCHK(F.info.ensure(3 * (size_t)d.nEL + 78));
  SolidInfoArgs a{};
  if (fill_solid_info_args(*b, body, refs, F.info.p, a)) return 1;
  hipLaunchKernelGGL(k_solid_info, dim3(1), dim3(256), 0, b->stream, a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rows, a.rows, 78 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_solid_info_static(int h, int body, const double *refs, double *rows) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  if (body < 0 || body >= F.nbodies) return fail("body %d out of range", body);
  if (!refs || !rows) return fail("null refs or rows");
  const auto it = g_scratch.find(g_device);
  if (it == g_scratch.end()) return fail("fsil_init has not set up the static scratch of device %d", g_device);
  const FemBodyDesc &d = F.desc[body];
  if (d.nEL > kStaticInfoElems)
    return fail("fsil_solid_info_static: body %d has %d elements, at most %d fit the static scratch (use "
                "fsil_solid_info)", body + 1, d.nEL, kStaticInfoElems);
  if (fem_collect(*b)) return 1;
  if (ibm_check_report(*b)) return 1;
  SolidInfoArgs a{};
  if (fill_solid_info_args(*b, body, refs, it->second, a)) return 1;
  hipLaunchKernelGGL(k_solid_info, dim3(1), dim3(256), 0, b->stream, a);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rows, a.rows, 78 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_fem_info(int h, int body, int *newton_iters, int *cg_iters, double *dnorm) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  if (body < 0 || body >= F.nbodies) return fail("body %d out of range", body);
  if (fem_collect(*b)) return 1;
  if (newton_iters) *newton_iters = F.desc[body].newton_iters;
  if (cg_iters) *cg_iters = F.desc[body].cg_iters;
  if (dnorm) *dnorm = F.desc[body].dnorm;
  return 0;
}

namespace {
// Beam-state checkpoint of a block's bodies: a header, the element records,
// then the node vectors in the order of Beam_write_solid_temp
// (SolidSolver.f90:1669-1677): pos, dsp, vel, acc; lodInte, lodExte,
// lodEffe; lodFlow, vBC.
constexpr char kCkpMagic[8] = {'F', 'S', 'I', 'L', 'C', 'K', 'P', '1'};
struct CkpHeader {
  char magic[8];
  int nbodies, nEL, nND, gEQ;
};
std::vector<std::pair<double *, size_t>> ckp_vectors(Block &b) {
  Fem &F = b.fem;
  const size_t g = 6 * (size_t)F.nND;
  return {{F.pos.p, g}, {F.dsp.p, g}, {F.vel.p, g}, {F.acc.p, g}, {F.lodInte.p, g}, {F.lodExte.p, g},
          {F.lodEffe.p, g}, {b.ibm.lod.p, (size_t)b.ibm.ndof}, {F.vBC.p, g}};
}
size_t ckp_bytes(Block &b) {
  size_t n = sizeof(CkpHeader) + (size_t)b.fem.nEL * sizeof(BeamElem);
  for (auto &v : ckp_vectors(b)) n += v.second * sizeof(double);
  return n;
}
}  // namespace

int fsil_fem_checkpoint(int h, void *buf, long long *bytes) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  if (F.nbodies <= 0) return fail("no beam bodies on block %d", h);
  if (!bytes) return fail("null size");
  const size_t need = ckp_bytes(*b);
  if (!buf) {
    *bytes = (long long)need;
    return 0;
  }
  if (*bytes < (long long)need) return fail("checkpoint buffer of %lld bytes, %zu needed", *bytes, need);
  if (fem_collect(*b)) return 1;
  if (ibm_check_report(*b)) return 1;
  char *o = (char *)buf;
  CkpHeader hd{};
  memcpy(hd.magic, kCkpMagic, 8);
  hd.nbodies = F.nbodies; hd.nEL = F.nEL; hd.nND = F.nND; hd.gEQ = 6 * F.nND;
  memcpy(o, &hd, sizeof(hd));
  o += sizeof(hd);
  HIPCHK(hipMemcpyAsync(o, F.el.p, (size_t)F.nEL * sizeof(BeamElem), hipMemcpyDeviceToHost, b->stream));
  o += (size_t)F.nEL * sizeof(BeamElem);
  for (auto &v : ckp_vectors(*b)) {
    HIPCHK(hipMemcpyAsync(o, v.first, v.second * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    o += v.second * sizeof(double);
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  *bytes = (long long)need;
  return 0;
}

int fsil_fem_restore(int h, const void *buf, long long bytes) {
  Block *b = get_block(h);
  if (!b) return 1;
  Fem &F = b->fem;
  if (F.nbodies <= 0) return fail("no beam bodies on block %d (define them first)", h);
  if (!buf || bytes < (long long)sizeof(CkpHeader)) return fail("checkpoint too short");
  CkpHeader hd;
  memcpy(&hd, buf, sizeof(hd));
  if (memcmp(hd.magic, kCkpMagic, 8) != 0) return fail("not a beam checkpoint");
  if (hd.nbodies != F.nbodies || hd.nEL != F.nEL || hd.nND != F.nND)
    return fail("checkpoint of %d bodies / %d elements / %d nodes, block has %d / %d / %d", hd.nbodies, hd.nEL,
                hd.nND, F.nbodies, F.nEL, F.nND);
  if (bytes < (long long)ckp_bytes(*b)) return fail("checkpoint truncated");
  if (fem_collect(*b)) return 1;
  const char *o = (const char *)buf + sizeof(hd);
  HIPCHK(hipMemcpyAsync(F.el.p, o, (size_t)F.nEL * sizeof(BeamElem), hipMemcpyHostToDevice, b->stream));
  o += (size_t)F.nEL * sizeof(BeamElem);
  for (auto &v : ckp_vectors(*b)) {
    HIPCHK(hipMemcpyAsync(v.first, o, v.second * sizeof(double), hipMemcpyHostToDevice, b->stream));
    o += v.second * sizeof(double);
  }
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

int fsil_copy_bandwidth(long long nbytes, int reps, double *gbs) {
  if (g_device < 0) return fail("fsil_init has not been called");
  if (nbytes < 1024 || reps <= 0 || !gbs) return fail("bad arguments to fsil_copy_bandwidth");
  const long long n = nbytes / (long long)sizeof(double);
  DevBuf<double> a, b;
  HIPCHK(a.ensure(n));
  HIPCHK(b.ensure(n));
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  hipLaunchKernelGGL(k_zero, grid, blk, 0, 0, a.p, n);
  hipLaunchKernelGGL(k_copy, grid, blk, 0, 0, a.p, b.p, n);
  HIPCHK(hipGetLastError());
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, grid, blk, 0, 0, a.p, b.p, n);
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbs = 2.0 * (double)n * sizeof(double) * reps / ((double)ms * 1e-3) / 1e9;
  return 0;
}

int fsil_ibm_band_info(int h, int *nband, int *bbox) {
  Block *b = get_block(h);
  if (!b) return 1;
  Ibm &I = b->ibm;
  int c[kCtlWords] = {0};
  if (I.ctl.p) {
    HIPCHK(hipMemcpyAsync(c, I.ctl.p, kCtlWords * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  const bool have = I.ctl.p && c[kNband] > 0;
  if (nband) *nband = have ? c[kNband] : 0;
  if (bbox) {
    for (int k = 0; k < 3; ++k) {
      bbox[k] = have ? c[kBoxLo + k] : 0;
      bbox[3 + k] = have ? c[kBoxHi + k] - c[kBoxLo + k] + 1 : 0;
    }
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// block tree (LBMBlockComm.f90)
namespace {

void share_stream(Block &b, const std::shared_ptr<StreamHolder> &sh) {
  if (b.stream != sh->s) (void)hipStreamSynchronize(b.stream);
  b.sh = sh;
  b.stream = sh->s;
  for (auto &p : b.sons) {
    auto it = g_blocks.find(p->son);
    if (it != g_blocks.end()) share_stream(*it->second, sh);
  }
}

void face_extents(const Pair &p, int face, int &aF, int &bF, int &aS, int &bS) {
  const int axis = face >> 1;
  const int ia = axis == 0 ? 1 : 0, ib = axis == 2 ? 1 : 2;
  aF = p.dimF[ia]; bF = p.dimF[ib];
  aS = p.dimS[ia]; bS = p.dimS[ib];
}

TreeArgs tree_args(Block &F, Block &S, const Pair &p) {
  TreeArgs t{};
  t.ff = F.fA.p; t.fs = S.fA.p; t.planes = p.planes.p;
  t.fxd = F.xd; t.fyd = F.yd; t.fzd = F.zd; t.fN = F.N;
  t.sxd = S.xd; t.syd = S.yd; t.szd = S.zd; t.sN = S.N;
  for (int a = 0; a < 3; ++a) {
    t.f0[a] = p.f[2 * a] - 1;
    t.fi0[a] = p.fi[2 * a] - 1;
    t.si0[a] = p.si[2 * a] - 1;
    t.sicnt[a] = p.si[2 * a + 1] >= p.si[2 * a] ? (p.si[2 * a + 1] - p.si[2 * a]) / 2 + 1 : 0;
  }
  for (int face = 0; face < 6; ++face) {
    TreeFace &tf = t.face[face];
    tf.active = p.sds[face] != 0;
    tf.fixedF = p.f[face] - 1;
    tf.fixedS = p.s[face] - 1;
    tf.fixedFi = p.fi[face] - 1;
    tf.fixedSi = p.si[face] - 1;
    face_extents(p, face, tf.aF, tf.bF, tf.aS, tf.bS);
    tf.off = p.off[face];
  }
  return t;
}

int tree_son(const Block &F, int index, Block *&S, Pair *&p) {
  if (index < 0 || index >= (int)F.sons.size())
    return fail("block %d has no son %d (%d sons)", F.id, index, (int)F.sons.size());
  p = F.sons[index].get();
  S = get_block(p->son);
  return S ? 0 : 1;
}

int extract_layer(Block &F, int level) {
  for (auto &pp : F.sons) {
    Pair &p = *pp;
    Block *S = get_block(p.son);
    if (!S) return 1;
    TreeArgs t = tree_args(F, *S, p);
    int maxn = 0;
    for (int face = 0; face < 6; ++face)
      if (t.face[face].active) maxn = std::max(maxn, Q * t.face[face].aF * t.face[face].bF);
    if (maxn == 0) continue;
    hipLaunchKernelGGL(k_tree_extract, dim3((maxn + 255) / 256, 6), dim3(256), 0, F.stream, t, level);
    HIPCHK(hipGetLastError());
    p.extracted = level;
  }
  return 0;
}

int father_to_son(Block &F, Block &S, Pair &p, int n_timeStep) {
  if (p.extracted != 2) return fail("father planes of block %d are not extracted at both levels", F.id);
  TreeArgs t = tree_args(F, S, p);
  // tau_all is the block's tau at every node for collision models 1-3, so the
  // interpolated father tau is the father's tau exactly
  t.coeff = (S.tau / F.tau) / 2.0;
  for (int k = 0; k < 3; ++k) t.hv[k] = 0.5 * F.VF[k] * F.dh;
  int maxn = 0;
  for (int face = 0; face < 6; ++face)
    if (t.face[face].active) maxn = std::max(maxn, t.face[face].aS * t.face[face].bS);
  if (maxn == 0) return 0;
  const dim3 grid((maxn + 255) / 256, 6);
  const int level = n_timeStep == 0 ? 1 : 2;
  if (p.scheme == 2) hipLaunchKernelGGL(k_tree_father_to_son<2>, grid, dim3(256), 0, F.stream, t, level);
  else hipLaunchKernelGGL(k_tree_father_to_son<1>, grid, dim3(256), 0, F.stream, t, level);
  HIPCHK(hipGetLastError());
  return 0;
}

int son_to_father(Block &F, Block &S, Pair &p) {
  TreeArgs t = tree_args(F, S, p);
  t.coeff = (F.tau / S.tau) * 2.0;
  for (int k = 0; k < 3; ++k) t.hv[k] = 0.5 * S.VF[k] * S.dh;
  int maxn = 0;
  for (int face = 0; face < 6; ++face) {
    if (!t.face[face].active) continue;
    const int axis = face >> 1;
    const int ea = axis == 0 ? 1 : 0, eb = axis == 2 ? 1 : 2;
    maxn = std::max(maxn, t.sicnt[ea] * t.sicnt[eb]);
  }
  if (maxn == 0) return 0;
  hipLaunchKernelGGL(k_tree_son_to_father, dim3((maxn + 255) / 256, 6), dim3(256), 0, F.stream, t);
  HIPCHK(hipGetLastError());
  return 0;
}

void set_tree_time(Block &b, double t) {
  b.blktime = t;
  for (auto &p : b.sons) {
    auto it = g_blocks.find(p->son);
    if (it != g_blocks.end()) set_tree_time(*it->second, t);
  }
}

// tree_collision_streaming_IBM_FEM for blocks without bodies
int tree_step(Block &b) {
  if (b.ibm.nbodies > 0 || b.fem.nbodies > 0)
    return fail("block %d carries bodies: step it from the host with the IBM/FEM calls", b.id);
  update_vf(b);
  if (extract_layer(b, 1)) return 1;
  if (collide_stream(b)) return 1;
  if (launch_faces(b)) return 1;
  b.ibm.active = false;
  if (extract_layer(b, 2)) return 1;
  for (auto &pp : b.sons) {
    Block *S = get_block(pp->son);
    if (!S) return 1;
    for (int n = 0; n < 2; ++n) {
      S->blktime = S->blktime + (double)n * S->dh;
      if (tree_step(*S)) return 1;
      if (father_to_son(b, *S, *pp, n)) return 1;
    }
    if (son_to_father(b, *S, *pp)) return 1;
  }
  return 0;
}

}  // namespace

extern "C" {

int fsil_tree_add_son(int father, int son, int interpolateScheme, int *index) {
  Block *F = get_block(father);
  if (!F) return 1;
  Block *S = get_block(son);
  if (!S) return 1;
  if (father == son) return fail("a block cannot be its own son");
  if (S->father) return fail("block %d already has a father (%d)", son, S->father);
  for (int a = father; a;) {
    if (a == son) return fail("block %d is an ancestor of block %d", son, father);
    auto it = g_blocks.find(a);
    a = it == g_blocks.end() ? 0 : it->second->father;
  }
  if (F->precision != 0 || S->precision != 0) return fail("block trees use the FP64 plain form");
  if (is_les(F->model) || is_les(S->model))
    return fail("block trees take collision models 1-3 (the LES models' per-cell tau_all is not exchanged)");
  // check_blocks_params (LBMBlockComm.f90:508-544)
  int r[3], per[3];
  for (int a = 0; a < 3; ++a) {
    per[a] = S->bc[2 * a] == BC_PERIODIC;
    r[a] = per[a] ? 0 : 1;
  }
  const bool flag = std::fabs(F->dh - S->dh * 2.0) > 1e-8 || S->xd % 2 != r[0] || S->yd % 2 != r[1] ||
                    S->zd % 2 != r[2];
  double res1 = (S->xmin - F->xmin) / F->dh + (S->ymin - F->ymin) / F->dh + (S->zmin - F->zmin) / F->dh;
  double res2 = (S->xmax - F->xmin) / F->dh + (S->ymax - F->ymin) / F->dh + (S->zmax - F->zmin) / F->dh;
  res1 = std::fabs(res1 - std::nearbyint(res1));
  res2 = std::fabs(res2 - std::nearbyint(res2));
  if (flag || res1 + res2 > 1e-8)
    return fail("grid points do not match between fluid blocks %d %d (%g %g); a son block needs an even number "
                "of grid points along periodic axes and an odd number otherwise",
                F->id, S->id, res1, res2);
  // build_blocks_comunication (LBMBlockComm.f90:47-95)
  auto pp = std::make_unique<Pair>();
  Pair &p = *pp;
  p.son = son;
  p.scheme = interpolateScheme == 2 ? 2 : 1;
  const int sd[3] = {S->xd - per[0], S->yd - per[1], S->zd - per[2]};
  const double smin[3] = {S->xmin, S->ymin, S->zmin}, fmin[3] = {F->xmin, F->ymin, F->zmin};
  const int ratio = (int)std::floor(F->dh / S->dh + 0.5);
  for (int j = 0; j < 6; ++j) p.sds[j] = S->bc[j] == BC_FLUID ? ((j & 1) ? -1 : 1) : 0;
  for (int a = 0; a < 3; ++a) {
    p.s[2 * a] = 1;
    p.s[2 * a + 1] = sd[a];
    p.f[2 * a] = (int)std::floor((smin[a] - fmin[a]) / F->dh + 1.5);
    p.f[2 * a + 1] = p.f[2 * a] + (sd[a] - 1) / ratio;
  }
  for (int j = 0; j < 6; ++j) {
    p.si[j] = p.s[j] + p.sds[j] * ratio;
    p.fi[j] = p.f[j] + p.sds[j];
  }
  p.dimS[0] = S->xd; p.dimS[1] = S->yd; p.dimS[2] = S->zd;
  for (int a = 0; a < 3; ++a) p.dimF[a] = p.f[2 * a + 1] - p.f[2 * a] + 1;
  // every index the exchange touches must lie inside both blocks
  const int fdim[3] = {F->xd, F->yd, F->zd}, sdim[3] = {S->xd, S->yd, S->zd};
  for (int a = 0; a < 3; ++a) {
    if (p.f[2 * a] < 1 || p.f[2 * a + 1] > fdim[a])
      return fail("son block %d does not lie inside father block %d along axis %d", S->id, F->id, a + 1);
    const int cnt = p.si[2 * a + 1] >= p.si[2 * a] ? (p.si[2 * a + 1] - p.si[2 * a]) / 2 + 1 : 0;
    if (cnt > 0 && (p.si[2 * a] < 1 || p.si[2 * a + 1] > sdim[a] || p.fi[2 * a] < 1 ||
                    p.fi[2 * a] + cnt - 1 > fdim[a]))
      return fail("son block %d: inner layers outside the blocks along axis %d", S->id, a + 1);
  }
  long long total = 0;
  for (int face = 0; face < 6; ++face) {
    if (!p.sds[face]) continue;
    const int a = face >> 1;
    if (p.s[face] < 1 || p.s[face] > sdim[a] || p.si[face] < 1 || p.si[face] > sdim[a] || p.f[face] < 1 ||
        p.f[face] > fdim[a] || p.fi[face] < 1 || p.fi[face] > fdim[a])
      return fail("son block %d: face %d layers outside the blocks", S->id, face + 1);
    int aF, bF, aS, bS;
    face_extents(p, face, aF, bF, aS, bS);
    const int aT = (aS & 1) ? aS : aS - 1, bT = (bS & 1) ? bS : bS - 1;
    if (p.scheme == 2 && (aT < 5 || bT < 5))
      return fail("4th-order interpolation needs son faces of at least 5 nodes (face %d is %d x %d)", face + 1, aS,
                  bS);
    if ((aT + 1) / 2 > aF || (bT + 1) / 2 > bF)
      return fail("son face %d does not match the father plane", face + 1);
    p.off[face] = total;
    total += 2LL * Q * aF * bF;
  }
  if (total > 0) HIPCHK(p.planes.ensure((size_t)total));
  share_stream(*S, F->sh);
  S->father = father;
  F->sons.push_back(std::move(pp));
  if (index) *index = (int)F->sons.size() - 1;
  return 0;
}

int fsil_tree_pair_info(int father, int index, int *comm) {
  Block *F = get_block(father);
  if (!F) return 1;
  Block *S;
  Pair *p;
  if (tree_son(*F, index, S, p)) return 1;
  if (!comm) return fail("null comm");
  for (int j = 0; j < 6; ++j) {
    comm[j] = p->sds[j];
    comm[6 + j] = p->s[j];
    comm[12 + j] = p->f[j];
    comm[18 + j] = p->si[j];
    comm[24 + j] = p->fi[j];
  }
  for (int a = 0; a < 3; ++a) {
    comm[30 + a] = p->dimS[a];
    comm[33 + a] = p->dimF[a];
  }
  return 0;
}

int fsil_extract_interpolate_layer(int father, int level) {
  Block *F = get_block(father);
  if (!F) return 1;
  if (level != 1 && level != 2) return fail("extract level must be 1 (old time) or 2 (new time)");
  return extract_layer(*F, level);
}

int fsil_interpolation_father_to_son(int father, int index, int n_timeStep) {
  Block *F = get_block(father);
  if (!F) return 1;
  Block *S;
  Pair *p;
  if (tree_son(*F, index, S, p)) return 1;
  if (n_timeStep != 0 && n_timeStep != 1) return fail("n_timeStep must be 0 or 1");
  return father_to_son(*F, *S, *p, n_timeStep);
}

int fsil_deliver_son_to_father(int father, int index) {
  Block *F = get_block(father);
  if (!F) return 1;
  Block *S;
  Pair *p;
  if (tree_son(*F, index, S, p)) return 1;
  return son_to_father(*F, *S, *p);
}

int fsil_tree_steps(int root, int nsteps, double t0, double dt) {
  Block *b = get_block(root);
  if (!b) return 1;
  if (b->father) return fail("block %d is not a tree root", root);
  double t = t0;
  for (int s = 0; s < nsteps; ++s) {
    t = t + dt;
    set_tree_time(*b, t);  // main.f90:97
    if (tree_step(*b)) return 1;
  }
  return 0;
}

int fsil_time_tree_steps(int root, int nsteps, double t0, double dt, double *total_ms) {
  Block *b = get_block(root);
  if (!b) return 1;
  if (nsteps <= 0 || !total_ms) return fail("bad arguments to fsil_time_tree_steps");
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  if (fsil_tree_steps(root, nsteps, t0, dt)) return 1;
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  *total_ms = ms;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// learned warm start: the network (fsil_mlp_kernels.h)
namespace {

uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + k * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double splitmix64_unit(uint64_t seed, uint64_t k) { return (double)(splitmix64_at(seed, k) >> 11) * (1.0 / 9007199254740992.0); }

int mlp_workspace(Mlp &M, int n) {
  HIPCHK(M.H1.ensure((size_t)n * kMlpH));
  HIPCHK(M.H2.ensure((size_t)n * kMlpH));
  HIPCHK(M.H3.ensure((size_t)n * kMlpH));
  HIPCHK(M.dZ3.ensure((size_t)n * kMlpH));
  HIPCHK(M.dZ2.ensure((size_t)n * kMlpH));
  HIPCHK(M.dZ1.ensure((size_t)n * kMlpH));
  HIPCHK(M.Yp.ensure((size_t)n * kMlpOut));
  HIPCHK(M.dY.ensure((size_t)n * kMlpOut));
  HIPCHK(M.part.ensure(((size_t)n * kMlpOut + 255) / 256));
  return 0;
}

inline dim3 g256(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// FP32 forward of n samples into M.H1..H3, M.Yp
int mlp_forward32(Mlp &M, const float *X, int n) {
  const float *th = M.th.p;
  hipLaunchKernelGGL(k_dense_fwd, g256((long long)n * kMlpH), dim3(256), 0, M.s, X, n, kMlpIn, th + kMlpOffW1,
                     th + kMlpOffB1, kMlpH, 1, M.H1.p);
  hipLaunchKernelGGL(k_dense_fwd, g256((long long)n * kMlpH), dim3(256), 0, M.s, M.H1.p, n, kMlpH, th + kMlpOffW2,
                     th + kMlpOffB2, kMlpH, 1, M.H2.p);
  hipLaunchKernelGGL(k_dense_fwd, g256((long long)n * kMlpH), dim3(256), 0, M.s, M.H2.p, n, kMlpH, th + kMlpOffW3,
                     th + kMlpOffB3, kMlpH, 1, M.H3.p);
  hipLaunchKernelGGL(k_dense_fwd, g256((long long)n * kMlpOut), dim3(256), 0, M.s, M.H3.p, n, kMlpH,
                     th + kMlpOffW4, th + kMlpOffB4, kMlpOut, 0, M.Yp.p);
  HIPCHK(hipGetLastError());
  return 0;
}

// forward + backward of n samples; gradient into M.g, squared-error block partials into M.part
int mlp_fwd_bwd(Mlp &M, const float *X, const float *Y, int n) {
  if (mlp_forward32(M, X, n)) return 1;
  const float *th = M.th.p;
  float *g = M.g.p;
  const int cnt = n * kMlpOut;
  hipLaunchKernelGGL(k_mse_grad, g256(cnt), dim3(256), 0, M.s, M.Yp.p, Y, cnt, M.dY.p, M.part.p);
  hipLaunchKernelGGL(k_dense_bwd_w, g256(kMlpOut * kMlpH + kMlpOut), dim3(256), 0, M.s, M.dY.p, M.H3.p, n, kMlpOut,
                     kMlpH, g + kMlpOffW4, g + kMlpOffB4);
  hipLaunchKernelGGL(k_dense_bwd_in, g256((long long)n * kMlpH), dim3(256), 0, M.s, M.dY.p, n, kMlpOut,
                     th + kMlpOffW4, kMlpH, M.H3.p, M.dZ3.p);
  hipLaunchKernelGGL(k_dense_bwd_w, g256(kMlpH * kMlpH + kMlpH), dim3(256), 0, M.s, M.dZ3.p, M.H2.p, n, kMlpH,
                     kMlpH, g + kMlpOffW3, g + kMlpOffB3);
  hipLaunchKernelGGL(k_dense_bwd_in, g256((long long)n * kMlpH), dim3(256), 0, M.s, M.dZ3.p, n, kMlpH,
                     th + kMlpOffW3, kMlpH, M.H2.p, M.dZ2.p);
  hipLaunchKernelGGL(k_dense_bwd_w, g256(kMlpH * kMlpH + kMlpH), dim3(256), 0, M.s, M.dZ2.p, M.H1.p, n, kMlpH,
                     kMlpH, g + kMlpOffW2, g + kMlpOffB2);
  hipLaunchKernelGGL(k_dense_bwd_in, g256((long long)n * kMlpH), dim3(256), 0, M.s, M.dZ2.p, n, kMlpH,
                     th + kMlpOffW2, kMlpH, M.H1.p, M.dZ1.p);
  hipLaunchKernelGGL(k_dense_bwd_w, g256(kMlpH * kMlpIn + kMlpH), dim3(256), 0, M.s, M.dZ1.p, X, n, kMlpH, kMlpIn,
                     g + kMlpOffW1, g + kMlpOffB1);
  HIPCHK(hipGetLastError());
  return 0;
}

int mlp_loss(Mlp &M, int n, double *loss) {
  const int cnt = n * kMlpOut, nb = (cnt + 255) / 256;
  std::vector<double> p(nb);
  HIPCHK(hipMemcpyAsync(p.data(), M.part.p, nb * sizeof(double), hipMemcpyDeviceToHost, M.s));
  HIPCHK(hipStreamSynchronize(M.s));
  double t = 0.0;
  for (double v : p) t = t + v;
  *loss = t / (double)cnt;
  return 0;
}

// rows idx[0..nb) of X/Y into the batch buffers
__global__ void k_gather_rows(const float *__restrict__ src, const int *__restrict__ idx, int nb, int width,
                              float *__restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * width) return;
  const int r = t / width, c = t - r * width;
  dst[t] = src[(size_t)idx[r] * width + c];
}

}  // namespace

extern "C" {

int fsil_mlp_create(int seed, int *handle) {
  if (g_device < 0) return fail("fsil_init has not been called");
  if (!handle) return fail("null handle");
  auto mp = std::make_unique<Mlp>();
  Mlp &M = *mp;
  HIPCHK(hipStreamCreateWithFlags(&M.s, hipStreamNonBlocking));
  std::vector<float> th(kMlpParams, 0.0f);
  const int sizes[4][2] = {{kMlpH, kMlpIn}, {kMlpH, kMlpH}, {kMlpH, kMlpH}, {kMlpOut, kMlpH}};
  int k = 0;
  for (auto &sz : sizes) {
    const double lim = std::sqrt(6.0 / (double)(sz[0] + sz[1]));
    for (int i = 0; i < sz[0] * sz[1]; ++i, ++k) th[k] = (float)(lim * (2.0 * splitmix64_unit((uint64_t)seed, k + 1) - 1.0));
    k += sz[0];  // zero biases
  }
  for (DevBuf<float> *b : {&M.th, &M.g, &M.m, &M.v}) HIPCHK(b->ensure(kMlpParams));
  HIPCHK(hipMemcpy(M.th.p, th.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(M.m.p, 0, kMlpParams * sizeof(float)));
  HIPCHK(hipMemset(M.v.p, 0, kMlpParams * sizeof(float)));
  const int h = g_next_mlp++;
  g_mlps[h] = std::move(mp);
  *handle = h;
  return 0;
}

int fsil_mlp_destroy(int handle) {
  if (!get_mlp(handle)) return 1;
  for (auto &kv : g_blocks)
    if (kv.second->ibm.warm_mlp == handle) kv.second->ibm.warm_mlp = 0;
  g_mlps.erase(handle);
  return 0;
}

int fsil_mlp_get_params(int handle, float *theta) {
  Mlp *M = get_mlp(handle);
  if (!M) return 1;
  if (!theta) return fail("null theta");
  HIPCHK(hipMemcpyAsync(theta, M->th.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, M->s));
  HIPCHK(hipStreamSynchronize(M->s));
  return 0;
}

int fsil_mlp_set_params(int handle, const float *theta) {
  Mlp *M = get_mlp(handle);
  if (!M) return 1;
  if (!theta) return fail("null theta");
  HIPCHK(hipMemcpyAsync(M->th.p, theta, kMlpParams * sizeof(float), hipMemcpyHostToDevice, M->s));
  HIPCHK(hipMemsetAsync(M->m.p, 0, kMlpParams * sizeof(float), M->s));
  HIPCHK(hipMemsetAsync(M->v.p, 0, kMlpParams * sizeof(float), M->s));
  HIPCHK(hipStreamSynchronize(M->s));
  M->adam_t = 0;
  return 0;
}

int fsil_mlp_forward(int handle, const float *X, int n, float *Y, int bf16) {
  Mlp *M = get_mlp(handle);
  if (!M) return 1;
  if (n <= 0 || !X || !Y) return fail("bad arguments to fsil_mlp_forward");
  HIPCHK(M->X.ensure((size_t)n * kMlpIn));
  HIPCHK(hipMemcpyAsync(M->X.p, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, M->s));
  if (mlp_workspace(*M, n)) return 1;
  if (bf16) {
    hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((n + 63) / 64), dim3(256), 0, M->s, M->X.p, n, M->th.p, M->Yp.p);
    HIPCHK(hipGetLastError());
  } else {
    if (mlp_forward32(*M, M->X.p, n)) return 1;
  }
  HIPCHK(hipMemcpyAsync(Y, M->Yp.p, (size_t)n * kMlpOut * sizeof(float), hipMemcpyDeviceToHost, M->s));
  HIPCHK(hipStreamSynchronize(M->s));
  return 0;
}

int fsil_mlp_loss_grad(int handle, const float *X, const float *Y, int n, double *loss, float *grad) {
  Mlp *M = get_mlp(handle);
  if (!M) return 1;
  if (n <= 0 || !X || !Y) return fail("bad arguments to fsil_mlp_loss_grad");
  HIPCHK(M->X.ensure((size_t)n * kMlpIn));
  HIPCHK(M->Y.ensure((size_t)n * kMlpOut));
  HIPCHK(hipMemcpyAsync(M->X.p, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, M->s));
  HIPCHK(hipMemcpyAsync(M->Y.p, Y, (size_t)n * kMlpOut * sizeof(float), hipMemcpyHostToDevice, M->s));
  if (mlp_workspace(*M, n)) return 1;
  if (mlp_fwd_bwd(*M, M->X.p, M->Y.p, n)) return 1;
  if (grad) HIPCHK(hipMemcpyAsync(grad, M->g.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, M->s));
  double l = 0.0;
  if (mlp_loss(*M, n, &l)) return 1;
  if (loss) *loss = l;
  return 0;
}

int fsil_mlp_train(int handle, const float *X, const float *Y, int n, int epochs, int batch, double lr, int seed,
                   double *loss_before, double *loss_after) {
  Mlp *M = get_mlp(handle);
  if (!M) return 1;
  if (n <= 0 || !X || !Y || epochs < 0 || batch <= 0) return fail("bad arguments to fsil_mlp_train");
  batch = std::min(batch, n);
  HIPCHK(M->X.ensure((size_t)n * kMlpIn));
  HIPCHK(M->Y.ensure((size_t)n * kMlpOut));
  HIPCHK(M->bx.ensure((size_t)batch * kMlpIn));
  HIPCHK(M->by.ensure((size_t)batch * kMlpOut));
  HIPCHK(M->idx.ensure((size_t)n));
  HIPCHK(hipMemcpyAsync(M->X.p, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, M->s));
  HIPCHK(hipMemcpyAsync(M->Y.p, Y, (size_t)n * kMlpOut * sizeof(float), hipMemcpyHostToDevice, M->s));
  if (mlp_workspace(*M, std::max(n, batch))) return 1;
  if (loss_before) {
    if (mlp_fwd_bwd(*M, M->X.p, M->Y.p, n)) return 1;
    if (mlp_loss(*M, n, loss_before)) return 1;
  }
  std::vector<int> perm(n);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) {  // Fisher-Yates from the splitmix64 stream
      const int j = (int)(splitmix64_at((uint64_t)seed + 7919ULL * (uint64_t)ep, (uint64_t)i) % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    HIPCHK(hipMemcpyAsync(M->idx.p, perm.data(), n * sizeof(int), hipMemcpyHostToDevice, M->s));
    HIPCHK(hipStreamSynchronize(M->s));  // perm is reused by the next epoch
    for (int b0 = 0; b0 + batch <= n; b0 += batch) {
      hipLaunchKernelGGL(k_gather_rows, g256((long long)batch * kMlpIn), dim3(256), 0, M->s, M->X.p, M->idx.p + b0,
                         batch, kMlpIn, M->bx.p);
      hipLaunchKernelGGL(k_gather_rows, g256((long long)batch * kMlpOut), dim3(256), 0, M->s, M->Y.p,
                         M->idx.p + b0, batch, kMlpOut, M->by.p);
      if (mlp_fwd_bwd(*M, M->bx.p, M->by.p, batch)) return 1;
      M->adam_t += 1;
      const float bc1 = 1.0f - std::pow(b1, (float)M->adam_t), bc2 = 1.0f - std::pow(b2, (float)M->adam_t);
      hipLaunchKernelGGL(k_adam, g256(kMlpParams), dim3(256), 0, M->s, M->th.p, M->g.p, M->m.p, M->v.p, kMlpParams,
                         (float)lr, b1, b2, eps, bc1, bc2);
      HIPCHK(hipGetLastError());
    }
  }
  if (loss_after) {
    if (mlp_fwd_bwd(*M, M->X.p, M->Y.p, n)) return 1;
    if (mlp_loss(*M, n, loss_after)) return 1;
  }
  HIPCHK(hipStreamSynchronize(M->s));
  return 0;
}

}  // extern "C"
