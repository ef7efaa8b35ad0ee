// ifdda_internal.h -- host-side plan and launcher declarations shared by the
// translation units of libifdda_hip.so (not part of the C ABI).
#pragma once

#include <string>

#include "ifdda_common.h"

namespace ifdda {

// Tile shape of one FFT pass: B batch lines per workgroup, `threads` threads.
struct TileShape {
  int B;
  int threads;
  size_t lds;  // dynamic LDS bytes
};

// Geometry of one box and its padded FFT grid.
struct GridPlan {
  int nx, ny, nz;   // box cells
  int L[3];         // padded FFT lengths (2,3,5-smooth, >= 2n)
  int F[3];         // folded lengths L/2 + 1
  AxisPlan ax[3];
  // tile shapes per pass and precision (index 0 = FP64, 1 = FP32)
  TileShape p1[2], p2[2], p3[2], p4[2], p5[2];
  TileShape gx, gy, gz;  // Green full-grid passes (FP64, 1 component)
};

struct PassTimer;

// Device-side matvec operands of one precision.
template <typename V> struct MatvecDev {
  V* A;               // nz*ny*L1*3 scratch (after x pass)
  V* B;               // nz*L2*L1*3 scratch (after y pass)
  const V* G;         // folded Green tensor, F2*F3*F1*6
  const V* tw[3];     // twiddles exp(+2 pi i e/L) per axis
  const V* tws[3];    // stage-major twiddles per axis (stage_twiddles)
  const int* pos[3];  // in-place FFT positions per axis (axis_positions)
  const int32_t* tab; // box -> 1-based dipole, or null (dense box)
  const V* alpha;     // per dipole, see alpha_mode
  int alpha_mode;
  PassTimer* timer = nullptr;
  bool specialized = true;  // use compile-time kernels where built
  // batch of right-hand sides: nrhs vectors vstride elements apart, A and B
  // sized for nrhs, skip flags skip_stride ints apart
  int nrhs = 1;
  long long vstride = 0;
  int skip_stride = 0;
};

// HIP-event timing of individual matvec passes on their own stream.  Slot
// tag = prec_index * 5 + pass (pass 0..4 = P1..P5).  Events come from a pool
// created when timing is enabled; recording stops when the pool is used up.
// Only the passes whose bit is set in mask are bracketed, and of those only
// every every-th launch (each event pair adds a few microseconds between
// launches).
struct PassTimer {
  int cap = 0;
  int used = 0;
  int mask = 31;
  int every = 1;
  long long seen = 0;  // launches of masked passes so far
  bool open = false;
  hipEvent_t* ev = nullptr;  // 2 * cap events
  int* tag = nullptr;
  void begin(int t, hipStream_t s) {
    open = used < cap && ((mask >> (t % 5)) & 1) && (seen++ % every == 0);
    if (open) {
      tag[used] = t;
      (void)hipEventRecord(ev[2 * used], s);
    }
  }
  void end(hipStream_t s) {
    if (open) (void)hipEventRecord(ev[2 * used + 1], s), ++used;
    open = false;
  }
};

// Report the first error (HIP or argument) and keep its text.
void set_error(const std::string& msg);
bool hip_ok(hipError_t e, const char* what);

int smooth235_at_least(int n);
bool make_axis_plan(int L, AxisPlan* out);
// pos[n]: LDS position of signal index n for the in-place mixed-radix FFT.
void axis_positions(const AxisPlan& ax, int* pos);
// Stage-major twiddles: stage s (m_s = r_1...r_{s-1}) holds exp(+2 pi i j q /
// (m_s r_s)) for j < m_s, q = 1..r_s-1 at index m_s - 1 + j (r_s - 1) + q - 1;
// L - 1 values in all (the stage sizes telescope).
void stage_twiddles(const AxisPlan& ax, double* re, double* im);
// false when an axis is too long for the register-staged FFT tiles
bool make_tiles(GridPlan& g, size_t lds_limit);
// whether the z pass of this grid has a compile-time kernel (prec 64 or 32)
bool has_specialized_z(const GridPlan& g, int prec);

// y = b - (1/M) IFFT[G . FFT[scatter(alpha . x)]]  (all device pointers).
// skip: optional device flag; when it reads non-zero the passes do nothing.
// timer: optional per-pass event timing.
bool launch_matvec_f64(const GridPlan& g, const MatvecDev<double2>& m, const double2* x,
                       const double2* b, double2* y, hipStream_t s, const int* skip = nullptr,
                       PassTimer* timer = nullptr);
bool launch_matvec_f32(const GridPlan& g, const MatvecDev<float2>& m, const float2* x,
                       const float2* b, float2* y, hipStream_t s, const int* skip = nullptr,
                       PassTimer* timer = nullptr);

// Green tensor on the device: fills g64/g32 (folded, F2*F3*F1*6) using
// `full` (L1*L2*L3 FP64 scratch).  nquad > 0 integrates near cells.
bool green_build_device(const GridPlan& g, const double2* tw64[3], const int* pos[3], double d,
                        double k0, int nquad, double2* full, double2* g64, float2* g32,
                        hipStream_t s);
// Fold one caller-supplied FFT'd component already in `full` into g64/g32.
bool green_fold_component(const GridPlan& g, const double2* full, int comp, double2* g64,
                          float2* g32, hipStream_t s);

}  // namespace ifdda

// The context behind the opaque ifdda_ctx handle of ifdda.h.
struct ifdda_ctx {
  ifdda::GridPlan g{};
  hipStream_t stream = nullptr;
  int device = 0;
  double2* tw64[3] = {nullptr, nullptr, nullptr};
  float2* tw32[3] = {nullptr, nullptr, nullptr};
  double2* tws64[3] = {nullptr, nullptr, nullptr};  // stage-major twiddles per axis
  float2* tws32[3] = {nullptr, nullptr, nullptr};
  int* pos[3] = {nullptr, nullptr, nullptr};  // in-place FFT positions per axis
  double2* G64 = nullptr;  // folded Green tensor, F2*F3*F1*6
  float2* G32 = nullptr;
  bool green_ready = false;
  int32_t* tab = nullptr;  // device dipole map (null: dense box)
  long long nd = 0;        // dipoles
  bool have_dipoles = false;
  double2* al64 = nullptr;  // polarisability (scalar or 3x3 row-major per dipole)
  float2* al32 = nullptr;
  int alpha_mode = ifdda::kAlphaNone;
  void* scrA = nullptr;  // FFT scratch sized for FP64 and scr_rhs right-hand sides
  void* scrB = nullptr;
  int scr_rhs = 1;
  double2* hx = nullptr;  // device staging for host-array entry points
  double2* hb = nullptr;
  size_t hcap = 0;        // complex elements in hx / hb
  void* solver = nullptr; // Krylov workspace (ifdda_krylov.hip)
  void* ffbuf = nullptr;  // far-field / cross-section workspace (ifdda_farfield.hip)
  size_t ffcap = 0;
  int32_t* dipcell = nullptr;  // device: box cell of each dipole (null: dense box)
  int cellbox[6] = {0, 0, 0, 0, 0, 0};  // occupied cells: i, j, k min / max (0-based)
  void* mbuf = nullptr;   // microscope workspace (ifdda_microscope.hip)
  size_t mcap = 0;
  // provenance of the device Green tensor for the reference-shaped entries
  const void* green_host = nullptr;
  double green_sample[6][2 * 16] = {};
  uint64_t tab_hash = 0;
  bool tab_dense = false;
  ifdda::PassTimer timer;  // enabled by ifdda_timing_enable
  bool specialized = true;  // ifdda_use_specialized

  ifdda::MatvecDev<double2> mv64() const {
    ifdda::MatvecDev<double2> m;
    m.A = (double2*)scrA;
    m.B = (double2*)scrB;
    m.G = G64;
    for (int i = 0; i < 3; ++i) m.tw[i] = tw64[i], m.tws[i] = tws64[i], m.pos[i] = pos[i];
    m.tab = tab;
    m.alpha = al64;
    m.alpha_mode = alpha_mode;
    m.timer = timer.cap ? const_cast<ifdda::PassTimer*>(&timer) : nullptr;
    m.specialized = specialized;
    return m;
  }
  ifdda::MatvecDev<float2> mv32() const {
    ifdda::MatvecDev<float2> m;
    m.A = (float2*)scrA;
    m.B = (float2*)scrB;
    m.G = G32;
    for (int i = 0; i < 3; ++i) m.tw[i] = tw32[i], m.tws[i] = tws32[i], m.pos[i] = pos[i];
    m.tab = tab;
    m.alpha = al32;
    m.alpha_mode = alpha_mode;
    m.timer = timer.cap ? const_cast<ifdda::PassTimer*>(&timer) : nullptr;
    m.specialized = specialized;
    return m;
  }
};

namespace ifdda {
// Release the Krylov workspace (ifdda_krylov.hip).
void solver_free(ifdda_ctx* c);
// Grow the FFT scratch to hold R right-hand sides (ifdda_capi.hip).
int ensure_scratch(ifdda_ctx* c, int R);
}  // namespace ifdda
