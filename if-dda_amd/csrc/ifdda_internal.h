// ifdda_internal.h -- host-side plan and launcher declarations shared by the
// translation units of libifdda_hip.so (not part of the C ABI).
#pragma once

#include <chrono>
#include <string>
#include <vector>

#include "../../include/ifdda.h"
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
  TileShape p1[2], p2[2], p3[2], p4[2], p5[2];  // batched launches and packs
  TileShape one[2][5];  // launches of one right-hand side: P1..P5 per precision
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
  // rows[k] bit j: box line (j, k) holds a dipole; null: every line (dense
  // boxes, ny > 64, or a map without empty lines).  Built by row_masks.
  const unsigned long long* rows = nullptr;
  const V* alpha;     // per dipole, see alpha_mode
  int alpha_mode;
  PassTimer* timer = nullptr;
  bool specialized = true;  // use compile-time kernels where built
  bool zreg = false;        // P3 as the register-staged kernel where built
  int dpar = 0;             // G is d T / d r_(dpar-1) (odd along that axis), 0: T
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
// Plan of a box: padded lengths, radix plans and tile shapes; false (and the
// reason) when a length has no plan or tile.
bool make_grid_plan(int nx, int ny, int nz, size_t lds_limit, GridPlan* out, std::string* why);
// Line masks of a dipole map (MatvecDev::rows): out[k] bit j set when a cell
// (i, j, k) maps to a dipole, k < nz.  False (out untouched) when they would
// skip nothing: ny > 64, or every line holds a dipole.
bool row_masks(const int32_t* tab, int nx, int ny, int nz, std::vector<unsigned long long>& out);
// Tables of one axis (L values each): twiddles exp(+2 pi i e / L) in FP64 and
// FP32, stage-major twiddles (stage_twiddles), in-place FFT positions.
void axis_tables(const AxisPlan& ax, double2* t64, float2* t32, double2* s64, float2* s32,
                 int* pos);
// Tile shape of pass p (0..4 = P1..P5) of precision index pi (0 = FP64,
// 1 = FP32): width exactly B, or the widest that fits lds_limit when B = 0;
// B = 0 in the result when the width does not fit.  (The autotuner's widths
// are those of one-right-hand-side launches, GridPlan::one.)
TileShape pass_tile_shape(const GridPlan& g, int pi, int p, int B, size_t lds_limit);
// Widths of P1..P5 recorded by the autotuner for a box and precision (32 or
// 64); false when none.
bool tuned_widths(int nx, int ny, int nz, int prec, int B[5]);
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

struct KState;
// GPBICG inner products fused into the last pass of a matvec (gpbicg.f:174-181,
// :222-253), for launches of one context: P5 forms the products of the
// vector it gathers, y = b - v/M, as it writes it.  Every vector holds R
// right-hand sides vstride apart.
//   p5 = 1: <r0*, Ap>, Ap = y of the matvec
//   p5 = 2: <At,At>, <y,y>, <At,y>, <At,t>, <y,t>, At = y of the matvec,
//           t = b of the matvec
// The products are the unfused kernels' (conj where they conjugate), in
// FP64, one partial per P5 block: part[(rhs * nd + q) * pstride + block].
template <typename V> struct KFuse {
  int p5 = 0;
  const V* r0s = nullptr;  // p5 = 1, 3
  const V* y = nullptr;    // p5 = 2, 3
  double2* part = nullptr;
  int pstride = 0;
  // p1 = 1: P1 first forms p = r + beta (p - u) (beta of each RHS's solver
  // state), stores it over p (the matvec input x) and scatters alpha p
  int p1 = 0;
  V* p = nullptr;
  const V* r = nullptr;
  const V* u = nullptr;
  const KState* st = nullptr;
};
// Blocks of the P5 pass of a launch over R right-hand sides (the partials a
// fused P5 writes per inner product and RHS).
int matvec_p5_blocks(const GridPlan& g, int prec, int R);
// The register-staged z pass (ifdda_zreg.hip): whether it is built for
// (L3, B) at prec 64 / 32, and its launch over a z pass's argument block and
// grid (24 B threads).  False when not built or not launchable.
template <typename V> struct SArgs;
bool has_zreg(int L, int B, int prec);
bool launch_zreg(const SArgs<double2>& a, int L, int B, dim3 grid, hipStream_t s);
bool launch_zreg(const SArgs<float2>& a, int L, int B, dim3 grid, hipStream_t s);
// The y passes with compile-time length and width (ifdda_ystride.hip): built
// for (L2, B) at prec and direction sgn (+1 P2, -1 P4); launch with the
// stage-major twiddles in a.tws and lds bytes of LDS.
bool has_strided_ct(int L, int B, int prec, int sgn);
bool launch_strided_ct(const SArgs<double2>& a, int L, int B, int sgn, dim3 grid, size_t lds,
                       hipStream_t s);
bool launch_strided_ct(const SArgs<float2>& a, int L, int B, int sgn, dim3 grid, size_t lds,
                       hipStream_t s);
bool launch_matvec_fused_f64(const GridPlan& g, const MatvecDev<double2>& m, const double2* x,
                             const double2* b, double2* y, hipStream_t s, const int* skip,
                             const KFuse<double2>& f);
bool launch_matvec_fused_f32(const GridPlan& g, const MatvecDev<float2>& m, const float2* x,
                             const float2* b, float2* y, hipStream_t s, const int* skip,
                             const KFuse<float2>& f);

// Packed matvec over P independent problems (members): the per-member
// argument blocks and block offsets of the five passes live in one device
// blob whose layout (byte offsets) is described here.
struct PackLayout {
  int P = 0;
  size_t av[5] = {}, boff[5] = {};  // byte offsets in the blob
  size_t mb = 0;                    // byte offset of the blocks per member and pass (5 x P ints)
  int nblk[5] = {};                 // blocks of each packed launch
  size_t lds[5] = {};               // dynamic LDS of each packed launch
  int ng3 = 1;                      // Green items per thread of the z pass
  std::vector<int> mbh;             // host copy: blocks of member m in pass p at [p * P + m]
};
// Rewrite the five block-offset tables of an uploaded blob (dev) so that only
// members with st[m].done == 0 (or every member, all) get blocks; lay.nblk
// follows (host_done: the same flags on the host, or null with all).
bool pack_compact(PackLayout& lay, void* dev, const KState* st, const int* host_done, bool all,
                  hipStream_t s);
// Host image of the blob for members (g[i], m[i]) whose dipole vectors start
// at element offs[i]; returns its size.
size_t pack_blob_f64(const std::vector<const GridPlan*>& g,
                     const std::vector<MatvecDev<double2>>& m, const long long* offs,
                     PackLayout& lay, std::vector<unsigned char>& blob);
size_t pack_blob_f32(const std::vector<const GridPlan*>& g,
                     const std::vector<MatvecDev<float2>>& m, const long long* offs,
                     PackLayout& lay, std::vector<unsigned char>& blob);
// y_m = b_m - (1/M_m) IFFT[G_m . FFT[scatter(alpha_m . x_m)]] for every
// member m (vectors at the blob's offsets); dev is the uploaded blob;
// skip[m * skip_stride] != 0 drops member m.
bool launch_matvec_pack_f64(const PackLayout& lay, const void* dev, const double2* x,
                            const double2* b, double2* y, hipStream_t s, const int* skip,
                            int skip_stride);
bool launch_matvec_pack_f32(const PackLayout& lay, const void* dev, const float2* x,
                            const float2* b, float2* y, hipStream_t s, const int* skip,
                            int skip_stride);

// Green tensor on the device: fills g64/g32 (folded, F2*F3*F1*6) using
// `full` (L1*L2*L3 FP64 scratch).  nquad > 0 integrates near cells.
// Grow-only device buffers of the near-cell quadrature (nquad > 0), kept by
// the owner of the builds (a context or a pack workspace) between builds.
struct GreenScratch {
  double2* near = nullptr;
  size_t near_cap = 0;  // bytes
  double2* part = nullptr;
  size_t part_cap = 0;
  int* offs = nullptr;
  size_t offs_cap = 0;
};
void green_scratch_free(GreenScratch& gs);
// The folded Green tensor of grid g (FP64, and FP32 when g32 is set).
// ncomp = 1: one component at a time, full holds M = L1 L2 L3 values;
// ncomp = 6: the six together in five launches, full holds 6 M values.
bool green_build_device(const GridPlan& g, const double2* tw64[3], const int* pos[3], double d,
                        double k0, int nquad, double2* full, double2* g64, float2* g32,
                        GreenScratch* gs, hipStream_t s, int ncomp = 1);
// The ncomp green_build_device takes for g in a pack: 6 while the six full
// grids stay within 512 MB, else 1.
int green_build_comps(const GridPlan& g);
// Fold one caller-supplied FFT'd component already in `full` into g64/g32.
bool green_fold_component(const GridPlan& g, const double2* full, int comp, double2* g64,
                          float2* g32, hipStream_t s);
// The folded FFT of d T / d r_dd (dd = 0, 1, 2) for the optical force
// (derivativefield2.f:22-76), FP64 only; full: L-grid scratch.
bool green_dbuild_device(const GridPlan& g, const double2* tw64[3], const int* pos[3], double d,
                         double k0, int dd, double2* full, double2* g64, hipStream_t s);

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
  unsigned long long* rows = nullptr;  // device line masks of tab (MatvecDev::rows)
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
  double2* sstage = nullptr;  // host-array solves: E_inc, FF, FFloc (grow-only)
  size_t scap = 0;            // complex elements in sstage
  void* solver = nullptr; // Krylov workspace (ifdda_krylov.hip)
  void* ffbuf = nullptr;  // far-field / cross-section workspace (ifdda_farfield.hip)
  size_t ffcap = 0;
  int32_t* dipcell = nullptr;  // device: box cell of each dipole (null: dense box)
  int cellbox[6] = {0, 0, 0, 0, 0, 0};  // occupied cells: i, j, k min / max (0-based)
  void* mbuf = nullptr;   // microscope workspace (ifdda_microscope.hip)
  size_t mcap = 0;
  double2* gfull = nullptr;  // full-grid scratch of the Green builds (grow-only)
  size_t gfull_cap = 0;      // bytes
  ifdda::GreenScratch gscr;  // near-cell quadrature buffers
  void* xbuf = nullptr;      // device staging of the host-array far-field entries (grow-only)
  size_t xcap = 0;
  // provenance of the device Green tensor for the reference-shaped entries:
  // the six host arrays and a hash of every 61st value of each
  const void* green_host[6] = {};
  uint64_t green_sig[6] = {};
  uint64_t tab_hash = 0;
  bool tab_dense = false;
  double2* Gd64 = nullptr;  // folded d T / d x, d y, d z (force), 3 * F2*F3*F1*6
  double gd_d = 0.0, gd_k0 = 0.0;  // the lattice and wavenumber Gd64 was built for
  bool gd_ready = false;            // cleared with every new Green tensor
  size_t lds_limit = 65536;  // dynamic LDS per workgroup the tiles may use
  ifdda::PassTimer timer;  // enabled by ifdda_timing_enable
  bool specialized = true;  // ifdda_use_specialized
  // GPBICG inner products fused into the matvec ends (ifdda_fuse_krylov):
  // bit 0 FP64 solves, bit 1 FP32 (the inner solves of mixed precision)
  int fuse_mask = 2 | 8;  // ifdda_fuse_krylov
  int zreg_mask = 0;      // ifdda_zpass_registers: bit 0 FP64, bit 1 FP32

  ifdda::MatvecDev<double2> mv64() const {
    ifdda::MatvecDev<double2> m;
    m.A = (double2*)scrA;
    m.B = (double2*)scrB;
    m.G = G64;
    for (int i = 0; i < 3; ++i) m.tw[i] = tw64[i], m.tws[i] = tws64[i], m.pos[i] = pos[i];
    m.tab = tab;
    m.rows = rows;
    m.alpha = al64;
    m.alpha_mode = alpha_mode;
    m.timer = timer.cap ? const_cast<ifdda::PassTimer*>(&timer) : nullptr;
    m.specialized = specialized;
    m.zreg = zreg_mask & 1;
    return m;
  }
  ifdda::MatvecDev<float2> mv32() const {
    ifdda::MatvecDev<float2> m;
    m.A = (float2*)scrA;
    m.B = (float2*)scrB;
    m.G = G32;
    for (int i = 0; i < 3; ++i) m.tw[i] = tw32[i], m.tws[i] = tws32[i], m.pos[i] = pos[i];
    m.tab = tab;
    m.rows = rows;
    m.alpha = al32;
    m.alpha_mode = alpha_mode;
    m.timer = timer.cap ? const_cast<ifdda::PassTimer*>(&timer) : nullptr;
    m.specialized = specialized;
    m.zreg = (zreg_mask >> 1) & 1;
    return m;
  }
};

namespace ifdda {
// Release the Krylov workspace (ifdda_krylov.hip).
void solver_free(ifdda_ctx* c);
// The context's grow-only full-grid Green scratch of M complex values (null
// on a HIP failure); a build synchronises before it returns.
double2* green_full(ifdda_ctx* c, size_t M);

// The operator A = I - T alpha of a solve over R right-hand sides whose
// vectors lie n apart: one context (all RHS share box, Green tensor, dipoles
// and polarisability) or a pack of independent problems.
struct MvOp {
  virtual ~MvOp() = default;
  virtual bool mv64(const double2* x, const double2* b, double2* y, hipStream_t s, const int* skip,
                    int R, long long n, int skip_stride) const = 0;
  virtual bool mv32(const float2* x, const float2* b, float2* y, hipStream_t s, const int* skip,
                    int R, long long n, int skip_stride) const = 0;
  // FF = alpha . x for every right-hand side
  virtual bool alpha_apply(const double2* x, double2* FF, int R, long long n,
                           hipStream_t s) const = 0;
  // The transposed operator y = x - alpha^T T x (what inverserig.f:559-590
  // forms for QMRCLA); zero: an all-zero vector of the layout of x.
  virtual bool mvt64(const double2* x, const double2* zero, double2* y, hipStream_t s,
                     const int* skip, int R, long long n, int skip_stride) const;
  virtual bool mvt32(const float2* x, const float2* zero, float2* y, hipStream_t s,
                     const int* skip, int R, long long n, int skip_stride) const;
  // Launch only the blocks of the RHS that are still running (st[r].done ==
  // 0, device state; host_done: the same flags on the host) from now on, or
  // of every RHS again (all = true).  Operators whose launches cover every
  // RHS in any case do nothing.
  virtual bool compact(const KState* st, const int* host_done, int R, bool all,
                       hipStream_t s) const {
    (void)st; (void)host_done; (void)R; (void)all; (void)s;
    return true;
  }
  // The matvec with GPBICG inner products fused into its end (KFuse).  fuse_blocks:
  // P5 blocks per RHS of such a launch, 0 when this operator does not fuse
  // at precision prec (the solver then runs the separate vector kernels).
  virtual int fuse_blocks(int prec, int R) const { (void)prec; (void)R; return 0; }
  virtual bool fuse_p1() const { return false; }  // p update inside P1 as well
  // the look-ahead GPBICG loop (three more fused products, no g_k1) for prec
  virtual bool fuse_la(int prec) const { (void)prec; return false; }
  virtual bool mvf64(const double2* x, const double2* b, double2* y, hipStream_t s,
                     const int* skip, int R, long long n, int skip_stride,
                     const KFuse<double2>& f) const {
    (void)x; (void)b; (void)y; (void)s; (void)skip; (void)R; (void)n; (void)skip_stride; (void)f;
    return false;
  }
  virtual bool mvf32(const float2* x, const float2* b, float2* y, hipStream_t s,
                     const int* skip, int R, long long n, int skip_stride,
                     const KFuse<float2>& f) const {
    (void)x; (void)b; (void)y; (void)s; (void)skip; (void)R; (void)n; (void)skip_stride; (void)f;
    return false;
  }
};
// y = x + alpha^T y for one right-hand side of nd dipoles (y holding -T x:
// the transposed operator's last step), unless *skip is set.
template <typename V>
bool alpha_t_launch(const V* x, V* y, const V* al, int mode, long long nd, hipStream_t s,
                    const int* skip);
// FF = alpha . x for one problem of nd dipoles.
bool alpha_apply_one(const double2* x, double2* FF, const double2* al, int mode, long long nd,
                     hipStream_t s);

// Krylov workspace (ifdda_krylov.hip) for R right-hand sides in the ragged
// layout of packed problems: RHS r holds lens[r] elements at offs[r] (host
// arrays of R values).
struct Solver;
Solver* solver_new(int R, bool need32, const long long* lens, const long long* offs,
                   long long min_total = 0, int min_R = 0);
// Capacity of S's buffers (zeros for a null S).
void solver_capacity(const Solver* S, long long* total, int* R);
// Reuse S for another ragged layout (lengths and offsets copied on stream s,
// the host arrays must live until s has run them); false when it is too small.
bool solver_refit(Solver* S, int R, bool need32, const long long* lens, const long long* offs,
                  hipStream_t s);
void solver_delete(Solver* S);
// The solve of R right-hand sides on S (FP64, or FP32 inner solves with FP64
// refinement): E_inc, FF, FFloc are device complex128 arrays in S's layout
// (FF / FFloc may be null).
int solve_core(const MvOp& op, Solver* S, const ifdda_solve_opts* o, int R, const double2* Einc,
               double2* FF, double2* FFloc, ifdda_solve_info* infos, volatile int* nstop,
               hipStream_t s, std::chrono::steady_clock::time_point t0);

// Incident waves and polarisabilities (ifdda_beams.hip).
// Polarisability of a cell of side d (poladiff, dddis = 1), IFDDA_POLA_*.
double2 poladiff_host(double2 eps, double d, double k0, double eps0, int method);
// ondeplanekxky.f polarisation vector (the wave's value at the origin) and kz.
void plane_wave_vector(double k0, double2 E0, double pp, double kx, double ky, double2 e[3],
                       double* kz_out);
// Wave vector k and polarisation e of one IFDDA_WAVE_* wave with parameters
// prm[3]; a non-empty message for invalid parameters.
std::string wave_vectors(int kind, const double* prm, double2 E0, double k0, double k[3],
                         double2 e[3]);
// E_r = e_r exp(i k_r . x) at the dipoles' cell centres origin + d (i, j, k)
// (dipcell: cell of each dipole, null for a dense box); RHS r at off[r] (or
// r * 3 nd); kv, ev: device arrays of R (k, e) triples.
bool launch_waves(double2* E, const long long* off, int R, long long nd, const int32_t* dipcell,
                  int nx, int ny, const double origin[3], double d, const double* kv,
                  const double2* ev, hipStream_t s);
// al64 / al32 = poladiff(eps) per dipole on the device.
bool launch_pola(const double2* eps, long long nd, double d, double k0, double eps0, int method,
                 double2* al64, float2* al32, hipStream_t s);

// Packs of independent problems (ifdda_pack.hip).  A workspace keeps its
// stream, FFT tables, Green tensors and grow-only buffers between packs.
struct PackWs;
PackWs* pack_ws_new();
void pack_ws_delete(PackWs* w);
// Bytes of Green tensors kept between packs (least recently used out).
void pack_ws_green_limit(PackWs* w, size_t bytes);
// One job of a pack: its host arrays and where its results go (FFloc / FF
// nrhs * 3 nd complex, may be null; infos nrhs entries).
struct PackIn {
  const ifdda_job* job;
  double* FFloc;
  double* FF;
  ifdda_solve_info* infos;
};
// Solves every right-hand side of every job together (one set of options).
int pack_solve(PackWs& w, const std::vector<PackIn>& in, const ifdda_solve_opts* o,
               volatile int* nstop);
// Device bytes a job adds to a pack (FFT scratch, vectors) and of its Green
// tensor.
size_t pack_member_bytes(const ifdda_job& j);
size_t pack_green_bytes(const ifdda_job& j);
// Grow the FFT scratch to hold R right-hand sides (ifdda_capi.hip).
int ensure_scratch(ifdda_ctx* c, int R);
// d E / d x_(dd) at every box cell of the field radiated by box-ordered
// moments (host, 3 N complex each; derivativefield2.f on the device).
int derivative_field_box(ifdda_ctx* c, const double* mom, double d, double k0, int dd, double* out);
}  // namespace ifdda
