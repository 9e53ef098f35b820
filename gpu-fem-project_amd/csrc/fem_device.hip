// HIP kernels of libfem_amd for gfx950 (MI355X, CDNA4, wave64) and their C-ABI launchers.
//
// Element math (LINEAR plane stress, reference element/FEKernels.h:229-315, closed forms of the A2D
// calls it makes):
//   J = sum_a x_a g_a^T, g_a = dN_a/dxi;  detJ signed;  gradN_a = J^-T g_a
//   grad u = sum_a u_a gradN_a^T;  eps = sym(grad u);  sigma = 2 mu eps + lambda tr(eps) I
//   s_q = w_q detJ_q t
//   R_{a,i}        = sum_q s_q sum_j sigma_ij gradN_{a,j}                         (FEKernels.h:106-197)
//   K_{(a,i),(b,k)} = sum_q s_q [lambda gN_{a,i} gN_{b,k} + mu (delta_ik gN_a.gN_b + gN_{a,k} gN_{b,i})]
// K is formed from the weighted Gram sums S_xx, S_yy, S_xy, S_yx of the physical gradients.
//
// Work mapping (all paths): a wavefront processes EPW = floor(64/N) elements at once; lane
// (el, idx) = (lane / N, lane % N). idx is a node in the gather, a quadrature point in the geometry
// stage and a column node b in the stiffness stage. The stiffness stage runs quadrature-point-outer:
// at each q every lane forms s_q*gradN_b(q), publishes it in a double-buffered per-wave LDS slot, and
// accumulates the Gram sums of its column b against every row a of a chunk of rows. Per-wave LDS
// therefore holds only the gathered nodes, the per-q geometry and one q's weighted gradients.
// Per-element LDS strides are padded to an odd number of 16-byte units so the broadcast reads of the
// elements of one wavefront fall in distinct banks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <type_traits>
#include <utility>
#include <cstring>
#include <vector>

#include "fem_internal.h"

namespace fem {

#define FEM_HIP_CHECK(expr)                                                                               \
  do {                                                                                                    \
    hipError_t err_ = (expr);                                                                             \
    if (err_ != hipSuccess) {                                                                             \
      return set_error(FEM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(err_), __FILE__,    \
                       __LINE__);                                                                         \
    }                                                                                                     \
  } while (0)

// Smallest count >= n of elements of `bytes` bytes spanning an odd number of 16-byte units.
constexpr int odd_units(int n, int bytes) {
  int s = n;
  while ((s * bytes) % 16 != 0 || ((s * bytes) / 16) % 2 == 0) s++;
  return s;
}

template <int P>
struct Ord {
  static constexpr int NP = P + 1;
  static constexpr int N = NP * NP;
  static constexpr int Q = N;
  static constexpr int EPW = 64 / N;
  static constexpr int LANES = EPW * N;
  static constexpr int RPAD = rank_pad(P);  // rank bytes stored per (element, column node)
  static constexpr int H = N / 2;
  static constexpr int D = H + 1;  // stiffness blocks per lane: rows a = b, b+1, ..., b+H (mod N)
  static constexpr int NODAL_STRIDE = 2 * odd_units(2 * N, 16);  // doubles; node a at [4a, 4a+4)
  static constexpr int GEO_STRIDE = 2 * odd_units(4 * Q, 16);    // doubles; point q at [8q, 8q+8)
};

// 1-D basis at the Gauss points: phi[k][i] = N_k(xi_i), dphi[k][i] = dN_k/dxi(xi_i), w[i].
struct Basis1D {
  double phi[5][5];
  double dphi[5][5];
  double w[5];
};

struct Material {
  double mu, lambda, t;
};

// Pointers read from device tables are generic to the compiler, which then issues flat loads; those also
// count against the LDS wait counter and would stall LDS waits on outstanding HBM loads. Data reached
// through these casts is loaded with global (vector-memory-only) instructions.
// (The host pass of the compiler parses device functions too; address spaces exist only in the device pass.)
#if defined(__HIP_DEVICE_COMPILE__)
#define FEM_GLOBAL_AS __attribute__((address_space(1)))
#else
#define FEM_GLOBAL_AS
#endif
template <typename T>
using gptr = const FEM_GLOBAL_AS T *;
template <typename T>
__device__ __forceinline__ gptr<T> gmem(const T *p) {
  return (gptr<T>)p;
}

template <typename T>
using gptr_mut = FEM_GLOBAL_AS T *;
// Read-only, wave-uniform tables (part descriptors) through the constant address space: scalar loads into
// scalar registers.
#if defined(__HIP_DEVICE_COMPILE__)
#define FEM_CONSTANT_AS __attribute__((address_space(4)))
#else
#define FEM_CONSTANT_AS
#endif
template <typename T>
__device__ __forceinline__ const FEM_CONSTANT_AS T *cmem(const T *p) {
  return (const FEM_CONSTANT_AS T *)p;
}
template <typename T>
__device__ __forceinline__ gptr_mut<T> gmem_mut(T *p) {
  return (gptr_mut<T>)p;
}

// Streaming (non-temporal) global accesses, for data touched once per assembly: the plan's visit tables and
// column ranks, and the BSR values written out. They are marked so that the L2 favours the node
// coordinates and states, which neighbouring parts gather again.
typedef double double2v __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ T load_once(gptr<T> p) {
  return __builtin_nontemporal_load(p);
}
__device__ __forceinline__ void store_once(gptr_mut<double2> p, double a, double b) {
  __builtin_nontemporal_store(double2v{a, b}, reinterpret_cast<FEM_GLOBAL_AS double2v *>(p));
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations, then s_barrier. Unlike
// __syncthreads() it does not wait for outstanding global loads and stores, so prefetches stay in flight
// across it. The memory clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Wave-uniform value (keeps part- and job-level indices in scalar registers).
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Distance in doubles between the four component planes of a part accumulator of MAXB blocks. The 4-double
// pad puts planes 0 and 2 (and 1 and 3) 64 bytes apart modulo 128, so the write-out's paired reads of
// block halves (xx, xy from planes 0-1 by even lanes, yx, yy from planes 2-3 by odd lanes) fall in
// different banks.
template <int MAXB>
constexpr int acc_plane() {
  static_assert(MAXB % 8 == 0, "plane offsets stay 16-byte aligned and 64 bytes apart modulo 128");
  return MAXB + 4;
}

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

// Per-wave LDS staging. The gathered nodes are dead once the geometry is formed, so they share
// storage with the weighted gradients of the stiffness stage.
template <int P, typename T>
struct __align__(16) WaveLDS {
  using O = Ord<P>;
  using V2 = typename Vec2<T>::type;
  static constexpr int SG_STRIDE = odd_units(O::N, int(sizeof(V2)));
  union {
    double nodal[O::EPW][O::NODAL_STRIDE];  // x, y, u, v of each element node
    V2 sg[2][O::EPW][SG_STRIDE];            // s_q * gradN_b(q) of the current quadrature point (double-buffered)
    float2 ns[O::EPW][4][O::N];             // MFMA stage: gradN_b at four quadrature points
  };
  double geo[O::EPW][O::GEO_STRIDE];  // per q: J^-1 (row-major), s_q, s_q*sigma_xx, s_q*sigma_yy, s_q*sigma_xy
  int obase[O::EPW][O::N];            // owner path: LDS block base of node a's row in this part, or -1
  uint32_t rk[O::EPW][O::N][O::RPAD / 4];  // owner path: rk[el][b] byte a = slot of column b within row a
  unsigned char orow[O::EPW][O::N];        // owner path, orders >= 3: the element's owned nodes, compacted
};

// Column rank of node `col` within BSR row of node `row` for element `el` (owner path).
template <typename WL>
__device__ __forceinline__ int rank_of(const WL &W, int el, int row, int col) {
  return int(reinterpret_cast<const uint8_t *>(&W.rk[el][col][0])[row]);
}

__device__ __forceinline__ void wave_sync() { __syncwarp(); }

// Workgroup index remapped so that the workgroups of one XCD (hardware placement: workgroup b on XCD
// b % 8) take a contiguous range of indices: neighbouring elements, whose outputs share cache lines, then
// meet in one L2.
__device__ __forceinline__ int xcd_block() {
  const int nb = int(gridDim.x), q = nb / 8, r = nb % 8, x = int(blockIdx.x) % 8;
  return x * q + min(x, r) + int(blockIdx.x) / 8;
}

__device__ __forceinline__ void copy_basis(Basis1D &dst, const Basis1D &src) {
  const double *s = reinterpret_cast<const double *>(&src);
  double *d = reinterpret_cast<double *>(&dst);
  constexpr int n = int(sizeof(Basis1D) / sizeof(double));
  for (int i = threadIdx.x; i < n; i += blockDim.x) d[i] = s[i];
}

// Reference gradients g_b(q) = (dN_b/dxi, dN_b/deta) at every quadrature point, gref[q*N + b].
template <int P>
__device__ __forceinline__ void fill_gref(double2 *gref, const Basis1D &src) {
  constexpr int NP = Ord<P>::NP, N = Ord<P>::N, Q = Ord<P>::Q;
  for (int i = threadIdx.x; i < Q * N; i += blockDim.x) {
    const int q = i / N, b = i - q * N;
    const int qy = q / NP, qx = q - qy * NP, by = b / NP, bx = b - by * NP;
    gref[i] = make_double2(src.phi[by][qy] * src.dphi[bx][qx], src.phi[bx][qx] * src.dphi[by][qy]);
  }
}

template <typename WL>
__device__ __forceinline__ void store_nodal(WL &W, int el, int a, double2 xy, double2 uv) {
  *reinterpret_cast<double2 *>(&W.nodal[el][4 * a]) = xy;
  *reinterpret_cast<double2 *>(&W.nodal[el][4 * a + 2]) = uv;
}

// Geometry and stress at quadrature point q of element `el` (lane (el, q)).
template <int P, typename T, typename WL>
__device__ __forceinline__ void stage_geometry(WL &W, const Basis1D &B, const Material &m, int el, int q) {
  constexpr int NP = Ord<P>::NP;
  const int qx = q % NP, qy = q / NP;
  double px[NP], dpx[NP], py[NP], dpy[NP];
#pragma unroll
  for (int k = 0; k < NP; k++) {
    px[k] = B.phi[k][qx];
    dpx[k] = B.dphi[k][qx];
    py[k] = B.phi[k][qy];
    dpy[k] = B.dphi[k][qy];
  }
  double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
  double U00 = 0, U01 = 0, U10 = 0, U11 = 0;
#pragma unroll
  for (int ay = 0; ay < NP; ay++) {
#pragma unroll
    for (int ax = 0; ax < NP; ax++) {
      const int a = ay * NP + ax;
      const double gx = py[ay] * dpx[ax];
      const double gy = px[ax] * dpy[ay];
      const double2 xy = *reinterpret_cast<const double2 *>(&W.nodal[el][4 * a]);
      const double2 uv = *reinterpret_cast<const double2 *>(&W.nodal[el][4 * a + 2]);
      J00 = fma(xy.x, gx, J00);
      J01 = fma(xy.x, gy, J01);
      J10 = fma(xy.y, gx, J10);
      J11 = fma(xy.y, gy, J11);
      U00 = fma(uv.x, gx, U00);
      U01 = fma(uv.x, gy, U01);
      U10 = fma(uv.y, gx, U10);
      U11 = fma(uv.y, gy, U11);
    }
  }
  const double detJ = J00 * J11 - J01 * J10;
  const double inv = 1.0 / detJ;
  const double i00 = J11 * inv, i01 = -J01 * inv, i10 = -J10 * inv, i11 = J00 * inv;
  const double G00 = U00 * i00 + U01 * i10;
  const double G01 = U00 * i01 + U01 * i11;
  const double G10 = U10 * i00 + U11 * i10;
  const double G11 = U10 * i01 + U11 * i11;
  const double mu = m.mu, lam = m.lambda;
  const double tr = G00 + G11;
  const double sxx = 2 * mu * G00 + lam * tr;
  const double syy = 2 * mu * G11 + lam * tr;
  const double sxy = mu * (G01 + G10);
  const double s = (B.w[qx] * B.w[qy]) * detJ * m.t;
  double *g = &W.geo[el][8 * q];
  *reinterpret_cast<double2 *>(g + 0) = make_double2(i00, i01);
  *reinterpret_cast<double2 *>(g + 2) = make_double2(i10, i11);
  *reinterpret_cast<double2 *>(g + 4) = make_double2(s, s * sxx);
  *reinterpret_cast<double2 *>(g + 6) = make_double2(s * syy, s * sxy);
}

// Geometry of one quadrature point as seen by lane (el, b): reference gradient of node b, J^-1 (rows
// ia, ib), s_q and the weighted stresses s_q*sigma.
struct PointGeo {
  double2 g, ia, ib, sc, sd;
};

template <int P, typename T, typename WL>
__device__ __forceinline__ void load_point(const WL &W, const double2 *gref, int el, int b, int q, PointGeo &G) {
  const double *g = &W.geo[el][8 * q];
  G.g = gref[q * Ord<P>::N + b];
  G.ia = *reinterpret_cast<const double2 *>(g + 0);
  G.ib = *reinterpret_cast<const double2 *>(g + 2);
  G.sc = *reinterpret_cast<const double2 *>(g + 4);
  G.sd = *reinterpret_cast<const double2 *>(g + 6);
}

// Physical gradient n = J^-T g of node b at the point, its weighted copy v = s_q n, and the point's
// residual contribution (B^T sigma s_q) added to (Rx, Ry).
template <typename T, typename V2>
// The residual is accumulated in FP64 in both precision modes; in mixed mode the gradient is rounded to
// FP32 for the Gram sums only.
__device__ __forceinline__ void point_terms(const PointGeo &G, T &nx, T &ny, V2 &v, double &Rx, double &Ry) {
  const double nxd = G.g.x * G.ia.x + G.g.y * G.ib.x;
  const double nyd = G.g.x * G.ia.y + G.g.y * G.ib.y;
  Rx += G.sc.y * nxd + G.sd.y * nyd;
  Ry += G.sd.y * nxd + G.sd.x * nyd;
  nx = T(nxd);
  ny = T(nyd);
  const T s = T(G.sc.x);
  v.x = s * nx;
  v.y = s * ny;
}

// Residual entries of node b and (JAC) stiffness blocks for lane (el, b). K is symmetric, so each
// unordered node pair {a, b} is formed once: lane b takes rows a = b + d (mod N), d = 0..N/2 (for even
// N the pair at distance N/2 is taken by the lower lane only), and both K(a, b) and K(b, a) = K(a, b)^T
// are emitted when a != b. where(row, col) gives the destination of block K(row, col) (negative: not
// stored here); emit(at, k) adds block k, row-major [xx, xy, yx, yy] in FP64, at destination `at`. All
// destinations are looked up before the first add, so the adds (which share the in-order wait counters
// with the lookups) never wait on one another. Every lane of the wavefront must call this (it
// synchronises the wavefront); `valid` masks the lane's LDS writes and emits.
// Software pipeline over quadrature points: iteration q publishes point q+1's weighted gradients and
// issues their reads, then accumulates point q's Gram sums from values read one iteration earlier, so
// LDS latency overlaps the arithmetic. The two publication buffers alternate; a buffer is rewritten two
// iterations after it was read, and the LDS executes one wavefront's accesses in order.
template <int P, typename T, bool JAC, typename Where, typename Emit>
__device__ __forceinline__ void stage_stiffness(WaveLDS<P, T> &W, const double2 *gref, const Material &m, int elIn,
                                                int b, bool valid, double &Rx, double &Ry, Where &&where,
                                                Emit &&emit) {
  using O = Ord<P>;
  constexpr int N = O::N, Q = O::Q, H = O::H, D = O::D;
  // Lanes past the last element of the wavefront (el == EPW) compute on element EPW-1's data: their LDS
  // reads stay inside the arrays, and `valid` masks their writes and emits.
  const int el = min(elIn, O::EPW - 1);
  using V2 = typename Vec2<T>::type;
  const T mu = T(m.mu), lam = T(m.lambda), c1 = lam + 2 * mu;
  Rx = 0;
  Ry = 0;
  T S[D][4];
#pragma unroll
  for (int i = 0; i < D; i++) S[i][0] = S[i][1] = S[i][2] = S[i][3] = 0;
  int arow[D];
#pragma unroll
  for (int d = 0; d < D; d++) arow[d] = b + d < N ? b + d : b + d - N;
  const bool lastPair = (N % 2 == 1) || b < H;  // even N: distance-N/2 pair taken by the lower lane

  // Gram sums of one point from the rows' weighted gradients wc and this lane's gradient (nx, ny).
  auto accumulate = [&](const V2(&wc)[D], T nx, T ny) __attribute__((always_inline)) {
    if (JAC) {
#pragma unroll
      for (int d = 0; d < D; d++) {
        S[d][0] = fma(wc[d].x, nx, S[d][0]);
        S[d][1] = fma(wc[d].y, ny, S[d][1]);
        S[d][2] = fma(wc[d].x, ny, S[d][2]);
        S[d][3] = fma(wc[d].y, nx, S[d][3]);
      }
    }
  };
  // One pipeline step (q + 1 < Q): publish point q+1 from G (its geometry), issue the reads of its rows
  // into wn and of its geometry successor into G, then accumulate point q from (wc, nx, ny).
  auto step = [&](int q, const V2(&wc)[D], V2(&wn)[D], PointGeo &G, T &nx, T &ny) __attribute__((always_inline)) {
    T nx1, ny1;
    V2 v1;
    point_terms<T, V2>(G, nx1, ny1, v1, Rx, Ry);
    load_point<P, T>(W, gref, el, b, q + 2 < Q ? q + 2 : Q - 1, G);
    if (JAC) {
      const int buf = (q + 1) & 1;
      if (valid) W.sg[buf][el][b] = v1;
      wave_sync();
#pragma unroll
      for (int d = 1; d < D; d++) wn[d] = W.sg[buf][el][arow[d]];
      wn[0] = v1;  // the lane's own row (arow[0] == b): its value is in registers
    }
    accumulate(wc, nx, ny);
    nx = nx1;
    ny = ny1;
  };

  // point 0
  PointGeo G;
  load_point<P, T>(W, gref, el, b, 0, G);
  T nx, ny;
  V2 v;
  point_terms<T, V2>(G, nx, ny, v, Rx, Ry);
  load_point<P, T>(W, gref, el, b, Q > 1 ? 1 : 0, G);
  V2 wA[D], wB[D];
  if (JAC) {
    if (valid) W.sg[0][el][b] = v;
    wave_sync();
#pragma unroll
    for (int d = 1; d < D; d++) wA[d] = W.sg[0][el][arow[d]];
    wA[0] = v;
  }
  // Points 0..Q-2 publish their successor; unrolled by two so that the register sets alternate instead
  // of being copied, with the odd step and the last point peeled at compile time.
  constexpr int PAIRS = (Q - 1) / 2;
#pragma unroll 1
  for (int i = 0; i < PAIRS; i++) {
    step(2 * i, wA, wB, G, nx, ny);
    step(2 * i + 1, wB, wA, G, nx, ny);
  }
  if constexpr ((Q - 1) % 2 == 1) {
    step(Q - 2, wA, wB, G, nx, ny);
    accumulate(wB, nx, ny);
  } else {
    accumulate(wA, nx, ny);
  }
  if (JAC && valid) {
    int atA[D], atB[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
      atA[d] = where(arow[d], b);
      atB[d] = d != 0 ? where(b, arow[d]) : -1;
    }
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (d == H && !lastPair) continue;
      double k[4];
      k[0] = double(c1 * S[d][0] + mu * S[d][1]);
      k[1] = double(lam * S[d][2] + mu * S[d][3]);
      k[2] = double(lam * S[d][3] + mu * S[d][2]);
      k[3] = double(c1 * S[d][1] + mu * S[d][0]);
      emit(atA[d], k);
      if (d != 0) {
        const double kt[4] = {k[0], k[2], k[1], k[3]};
        emit(atB[d], kt);
      }
    }
  }
}

// ---- element-parallel kernel (drop-in atomic path and colour path) -----------------------------
template <int P, typename T, bool JAC, bool ALL_LANES, typename Where, typename Emit>
__device__ __forceinline__ void stage_stiffness_rows(WaveLDS<P, T> &W, const double2 *gref, const Material &m, int el,
                                                     int b, bool valid, int count, int maxCount, double &Rx,
                                                     double &Ry, Where &&where, Emit &&emit);

struct ElemArgs {
  const int *connPtr;
  const int *conn;
  const int *elemList;  // colour path: element ids of this colour; NULL: elements first..first+count-1
  int first;
  int count;
  const double *states;
  const double *coords;
  const int *map;
  double *residual;
  double *values;
  Basis1D basis;
  Material mat;
};

template <int P, typename T, int WAVES, bool RES, bool JAC, bool ATOMIC>
__global__ __launch_bounds__(WAVES * 64) void element_kernel(const ElemArgs args) {
  using O = Ord<P>;
  constexpr int N = O::N;
  __shared__ WaveLDS<P, T> wl[WAVES];
  __shared__ Basis1D basis;
  __shared__ double2 gref[O::Q * O::N];
  __shared__ double stK[JAC ? WAVES : 1][4][64];  // Jacobian emission: each lane's block, by component
  __shared__ int stAt[JAC ? WAVES : 1][64];        // and its destination (negative: none)
  copy_basis(basis, args.basis);
  fill_gref<P>(gref, args.basis);
  __syncthreads();

  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int el = lane / N, idx = lane % N;
  WaveLDS<P, T> &W = wl[wave];
  const int64_t slot = (int64_t(xcd_block()) * WAVES + wave) * O::EPW + el;
  const bool valid = lane < O::LANES && slot < args.count;
  int eid = 0, node = 0;
  if (valid) {
    eid = args.elemList ? args.elemList[slot] : args.first + int(slot);
    node = args.conn[args.connPtr[eid] + idx];
    store_nodal(W, el, idx, *reinterpret_cast<const double2 *>(args.coords + 2 * int64_t(node)),
                *reinterpret_cast<const double2 *>(args.states + 2 * int64_t(node)));
  }
  wave_sync();
  if (valid) stage_geometry<P, T>(W, basis, args.mat, el, idx);
  wave_sync();
  const int *mel = args.map + int64_t(eid) * N * N;
  double Rx, Ry;
  auto where = [&](int row, int col) __attribute__((always_inline)) { return mel[row * N + col]; };
  auto emit = [&](int at, const double(&k)[4]) __attribute__((always_inline)) {
    double *blk = args.values + at;
    if (ATOMIC) {
#pragma unroll
      for (int c = 0; c < 4; c++) unsafeAtomicAdd(blk + c, k[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 4; c++) blk[c] += k[c];
    }
  };
  if constexpr (JAC) {
    // Row-major emission: every row of K_e in chunks, so the lanes of an element hold neighbouring blocks of
    // one BSR row. The blocks then pass through LDS so that each add instruction covers the four
    // components of 16 blocks (lanes 4i..4i+3 take block i), 32 contiguous bytes per block.
    if (valid) W.orow[el][idx] = (unsigned char)idx;
    wave_sync();
    auto emitRows = [&](int at, const double(&k)[4]) __attribute__((always_inline)) {
      stAt[wave][lane] = at;
#pragma unroll
      for (int c = 0; c < 4; c++) stK[wave][c][lane] = k[c];
      wave_sync();
      const int c = lane & 3;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int src = (lane >> 2) + 16 * r;
        const int a = stAt[wave][src];
        if (a >= 0) {
          double *blk = args.values + a + c;
          if (ATOMIC) unsafeAtomicAdd(blk, stK[wave][c][src]);
          else *blk += stK[wave][c][src];
        }
      }
      wave_sync();  // the next row restages into the same slots
    };
    stage_stiffness_rows<P, T, JAC, true>(W, gref, args.mat, el, idx, valid, N, N, Rx, Ry, where, emitRows);
  } else {
    stage_stiffness<P, T, JAC>(W, gref, args.mat, el, idx, valid, Rx, Ry, where, emit);
  }
  if (RES && valid) {
    double *r = args.residual + 2 * int64_t(node);
    if (ATOMIC) {
      unsafeAtomicAdd(r, double(Rx));
      unsafeAtomicAdd(r + 1, double(Ry));
    } else {
      r[0] += double(Rx);
      r[1] += double(Ry);
    }
  }
}

// ---- element-once assembly of a plan with split tables, orders 3-4 -----------------------------------
// K_e is formed in tiles of 4 x 4 nodes, one tile per lane, upper tiles only (K_e is symmetric): for G =
// ceil(N / 4) node groups, tile t of an element covers rows 4I..4I+3 and columns 4J..4J+3, I <= J, in the
// order (0,0), (0,1), .., (0,G-1), (1,1), ... Per quadrature point each lane (el, a) publishes node a's
// physical gradient n_a and its weighted copy s_q n_a (adding node a's residual terms), and each tile lane
// reads the weighted gradients of its 4 rows and the gradients of its 4 columns and adds the Gram sums of
// its 16 blocks: 64 FMAs for 8 LDS reads.
// Calls f(std::integral_constant<int, i>) for i = 0 .. n-1, each call a separate instantiation.
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int n, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(f, std::make_integer_sequence<int, n>{});
}

template <int P>
struct TileMap {
  static constexpr int N = Ord<P>::N;
  static constexpr int G = (N + 3) / 4;
  static constexpr int T = G * (G + 1) / 2;
  static_assert(Ord<P>::EPW * T <= 64, "the tiles of a wavefront's elements exceed its lanes");
};

// The emission's store instructions: group R of 4 rows takes iters(R) instructions (16 blocks each over
// the wavefront's elements), issued in chunks of CHK whose destination codes are loaded one chunk ahead
// (order 3: 16 instructions per group in two chunks, so the codes fit the registers beside the Gram tiles).
template <int P>
struct EmitPlan {
  static constexpr int N = Ord<P>::N, EPW = Ord<P>::EPW, G = TileMap<P>::G;
  static constexpr int CHK = P == 3 ? 8 : 16;
  static constexpr int iters(int R) { return (EPW * (N - 4 * R < 4 ? N - 4 * R : 4) * N + 15) / 16; }
  static constexpr int chunks(int R) { return (iters(R) + CHK - 1) / CHK; }
  static constexpr int steps() {
    int t = 0;
    for (int R = 0; R < G; R++) t += chunks(R);
    return t;
  }
  static constexpr int group(int k) {
    for (int R = 0; R < G; R++) {
      if (k < chunks(R)) return R;
      k -= chunks(R);
    }
    return G - 1;
  }
  static constexpr int chunk(int k) {
    for (int R = 0; R < G; R++) {
      if (k < chunks(R)) return k;
      k -= chunks(R);
    }
    return 0;
  }
};

template <int P, typename T>
struct __align__(16) WaveLDSTile {
  using O = Ord<P>;
  using V2 = typename Vec2<T>::type;
  static constexpr int PUB_STRIDE = odd_units(O::N, int(sizeof(V2)));
  // Emission staging of one group of 4 rows: plane 0 holds (K_xx, K_xy) and plane 1 (K_yx, K_yy) of block
  // (el, row i, column b) at kidx(el, i, b). Column b sits at b + b/4 and elements are an odd number of
  // 2-vectors apart, so the lanes of one staging write (one per tile: columns 4 apart) hit distinct banks.
  static constexpr int KROW = O::N + O::N / 4 + 1;
  static constexpr int KEL = 4 * KROW + 1;
  union {
    double nodal[O::EPW][O::NODAL_STRIDE];  // x, y, u, v of each element node
    V2 pub[2][2][O::EPW][PUB_STRIDE];       // [buffer][n_a, s_q n_a][el][a] at the current point
    V2 kq[2][O::EPW * KEL];
  };
  double geo[O::EPW][O::GEO_STRIDE];  // as WaveLDS::geo

  static __device__ __forceinline__ int kidx(int el, int i, int b) { return el * KEL + i * KROW + b + b / 4; }
  __device__ __forceinline__ void put(int el, int i, int b, T k0, T k1, T k2, T k3) {
    V2 lo, hi;
    lo.x = k0;
    lo.y = k1;
    hi.x = k2;
    hi.y = k3;
    kq[0][kidx(el, i, b)] = lo;
    kq[1][kidx(el, i, b)] = hi;
  }
  __device__ __forceinline__ T get(int el, int i, int b, int c) const {
    return reinterpret_cast<const T *>(&kq[c >> 1][kidx(el, i, b)])[c & 1];
  }
};

// One element-once assembly: a plan's element-once tables and a request's inputs and outputs.
struct SplitJob {
  const int *conn;   // [E][N]
  const int *code;   // [E][N][N] block codes (SplitHost::code)
  const int *rcode;  // [E][N] residual codes
  const int *gdst, *gptr, *rdst, *rptr;  // gather lists of the staged blocks and residual entries
  const double *states, *coords;
  double *residual;  // NULL: no residual
  double *values;
  double *stage, *rstage;
  Material mat;
  int count, numBlocks, numNodes, numGather, numRgather;
  int accumulate;  // 1: outputs are added to (+=), staging is still written
};

// Elements first .. first+EPW-1 of `job` (those below job.count) on one wavefront: geometry, Gram tiles,
// residual and emission. Every lane of the wavefront calls it; it synchronises the wavefront only.
template <int P, typename T, bool RES, bool ACC = false>
__device__ __forceinline__ void tiled_wave(WaveLDSTile<P, T> &W, const Basis1D &basis, const double2 *gref,
                                           const SplitJob &job, int64_t first, int lane) {
  using O = Ord<P>;
  using TM = TileMap<P>;
  using V2 = typename Vec2<T>::type;
  constexpr int N = O::N, Q = O::Q, EPW = O::EPW, NT = TM::T;
  static_assert(Q == N, "lane (el, q) of the geometry stage is lane (el, a)");
  wave_sync();  // a previous call's LDS reads are done before the nodes overwrite them
  if (first >= job.count) return;  // no element: the emission below assumes at least one
  // node role: lane (el, a)
  const int el = lane / N, a = lane % N;
  const bool nvalid = lane < EPW * N && first + el < job.count;
  const int elc = el < EPW ? el : EPW - 1;
  int node = 0;
  if (nvalid) {
    node = job.conn[(first + el) * N + a];
    store_nodal(W, el, a, *reinterpret_cast<const double2 *>(job.coords + 2 * int64_t(node)),
                *reinterpret_cast<const double2 *>(job.states + 2 * int64_t(node)));
  }
  wave_sync();
  if (nvalid) stage_geometry<P, T>(W, basis, job.mat, el, a);
  wave_sync();
  // tile role: lane (te, t)
  const int te = lane / NT;
  const bool tvalid = lane < EPW * NT && first + te < job.count;
  const int tec = te < EPW ? te : EPW - 1;
  int I = 0, J = 0;
  {
    int r = lane < EPW * NT ? lane % NT : 0;
    while (r >= TM::G - I) {
      r -= TM::G - I;
      I++;
    }
    J = I + r;
  }
  int rows[4], cols[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    rows[k] = min(4 * I + k, N - 1);  // past the last node: a real node, not emitted
    cols[k] = min(4 * J + k, N - 1);
  }
  T S[4][4][4];  // [row][col][xx, yy, xy, yx]
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) S[i][j][0] = S[i][j][1] = S[i][j][2] = S[i][j][3] = 0;
  double Rx = 0, Ry = 0;
#pragma unroll 1
  for (int q = 0; q < Q; q++) {
    const int buf = q & 1;
    PointGeo G;
    load_point<P, T>(W, gref, elc, a < N ? a : 0, q, G);
    T nx, ny;
    V2 v;
    point_terms<T, V2>(G, nx, ny, v, Rx, Ry);
    if (nvalid) {
      W.pub[buf][0][el][a].x = nx;
      W.pub[buf][0][el][a].y = ny;
      W.pub[buf][1][el][a] = v;
    }
    wave_sync();  // a buffer is rewritten two points after it was read; a wavefront's LDS accesses run in order
    V2 rw[4], cn[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      rw[k] = W.pub[buf][1][tec][rows[k]];
      cn[k] = W.pub[buf][0][tec][cols[k]];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        S[i][j][0] = fma(rw[i].x, cn[j].x, S[i][j][0]);
        S[i][j][1] = fma(rw[i].y, cn[j].y, S[i][j][1]);
        S[i][j][2] = fma(rw[i].x, cn[j].y, S[i][j][2]);
        S[i][j][3] = fma(rw[i].y, cn[j].x, S[i][j][3]);
      }
    }
  }
  if (RES && nvalid) {
    const int rc = job.rcode[(first + el) * N + a];
    double *r = rc < job.numNodes ? job.residual + 2 * int64_t(rc) : job.rstage + 2 * (int64_t(rc) - job.numNodes);
    if (ACC && rc < job.numNodes) {
      Rx += r[0];
      Ry += r[1];
    }
    __builtin_nontemporal_store(Rx, r);
    __builtin_nontemporal_store(Ry, r + 1);
  }
  // The tile's blocks K(a, b), row-major [xx, xy, yx, yy], in place of its Gram sums.
  {
    const T mu = T(job.mat.mu), lam = T(job.mat.lambda), c1 = lam + 2 * mu;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const T xx = S[i][j][0], yy = S[i][j][1], xy = S[i][j][2], yx = S[i][j][3];
        S[i][j][0] = c1 * xx + mu * yy;
        S[i][j][1] = lam * xy + mu * yx;
        S[i][j][2] = lam * yx + mu * xy;
        S[i][j][3] = c1 * yy + mu * xx;
      }
    }
  }
  // Emission, one group of 4 rows at a time: the tiles holding blocks of rows 4R..4R+3 write them to LDS
  // (tile (R, J) directly; tile (J, R), J < R, transposed: K(b, a) = K(a, b)^T), then each store instruction
  // covers the four components of 16 consecutive blocks of the group's rows (lanes 4k..4k+3: block k). The
  // destination codes of the next chunk of stores (EmitPlan) are loaded while the current chunk is stored.
  // Every code load and store is unconditional, so that the wait for a chunk's codes lets the previous
  // chunk's stores stay in flight:
  // lanes past the group's last block (or past the wavefront's last element) repeat the store of that block,
  // the same value to the same address in the same instruction.
  using EP = EmitPlan<P>;
  constexpr int CHK = EP::CHK;
  const int c = lane & 3;
  const int lastEl = int(min<int64_t>(EPW, job.count - first)) - 1;
  // block (e, rem) of store instruction `it` of group R (rem: offset within the element's rows of the group);
  // dup: the lane repeats another lane's block (past the group's last block or the wavefront's last element)
  auto locate = [&](int R, int it, int &e, int &rem, bool &dup) __attribute__((always_inline)) {
    const int groupBlocks = (N - 4 * R < 4 ? N - 4 * R : 4) * N;
    const int want = 16 * it + (lane >> 2);
    dup = want > EPW * groupBlocks - 1;
    rem = min(want, EPW * groupBlocks - 1);
    e = 0;
#pragma unroll
    for (int k = 1; k < EPW; k++) {
      if (rem >= groupBlocks) {
        rem -= groupBlocks;
        e++;
      }
    }
    if (e > lastEl) {
      e = lastEl;
      rem = groupBlocks - 1;
      dup = true;
    }
  };
  int cur[CHK], nxt[CHK];
  // codes of step k: chunk EP::chunk(k) of group EP::group(k)
  auto load_codes = [&](auto Kc, int (&at)[CHK]) __attribute__((always_inline)) {
    constexpr int k = decltype(Kc)::value, R = EP::group(k), ch = EP::chunk(k);
#pragma unroll
    for (int it = 0; it < CHK; it++) {
      if (ch * CHK + it >= EP::iters(R)) break;
      int e, rem;
      bool dup;
      locate(R, ch * CHK + it, e, rem, dup);
      at[it] = job.code[((first + e) * N + 4 * R) * N + rem];
    }
  };
  load_codes(std::integral_constant<int, 0>{}, cur);
  static_for<EP::steps()>([&](auto Kc) __attribute__((always_inline)) {
    constexpr int k = decltype(Kc)::value, R = EP::group(k), ch = EP::chunk(k);
    if constexpr (k + 1 < EP::steps()) load_codes(std::integral_constant<int, k + 1>{}, nxt);
    if constexpr (ch == 0) {
      wave_sync();  // the previous group's (or the last point's) LDS reads are done
      if (tvalid && I == R) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (4 * I + i < N && 4 * J + j < N) W.put(tec, i, 4 * J + j, S[i][j][0], S[i][j][1], S[i][j][2], S[i][j][3]);
      } else if (tvalid && J == R) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (4 * J + j < N) W.put(tec, j, 4 * I + i, S[i][j][0], S[i][j][2], S[i][j][1], S[i][j][3]);
      }
      wave_sync();
    }
#pragma unroll
    for (int it = 0; it < CHK; it++) {
      if (ch * CHK + it >= EP::iters(R)) break;
      int e, rem;
      bool dup;
      locate(R, ch * CHK + it, e, rem, dup);
      const int at = cur[it];
      const int i = rem / N, b = rem - i * N;
      double *blkp = at < job.numBlocks ? job.values + 4 * int64_t(at) : job.stage + 4 * (int64_t(at) - job.numBlocks);
      double val = double(W.get(e, i, b, c));
      if (!ACC) {
        __builtin_nontemporal_store(val, blkp + c);
      } else if (!dup) {
        // accumulate mode: each block added once (repeating lanes store nothing); staging is written
        if (at < job.numBlocks) val += blkp[c];
        __builtin_nontemporal_store(val, blkp + c);
      }
    }
#pragma unroll
    for (int it = 0; it < CHK; it++) cur[it] = nxt[it];
  });
}


// Two wavefronts per SIMD, which the LDS of two 4-wavefront workgroups per CU allows: without the bound the FP64
// order-3 kernel takes 259 registers and runs one wavefront per SIMD (order 3 FP64: 1.40 against 1.50 ms).
template <int P, typename T, int WAVES, bool RES, bool ACC = false>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(2)))
void element_tiled_kernel(const SplitJob job, const Basis1D basisIn) {
  using O = Ord<P>;
  __shared__ WaveLDSTile<P, T> wl[WAVES];
  __shared__ Basis1D basis;
  __shared__ double2 gref[O::Q * O::N];
  copy_basis(basis, basisIn);
  fill_gref<P>(gref, basisIn);
  __syncthreads();
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  tiled_wave<P, T, RES, ACC>(wl[wave], basis, gref, job, (int64_t(xcd_block()) * WAVES + wave) * O::EPW, lane);
}

// Index j of the last start[j] <= x over start[0..n) (start ascending, start[0] = 0 <= x).
__device__ __forceinline__ int find_job(const int *start, int n, int64_t x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (start[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A batch's element-once jobs of one order: counts = (jobs, units, staged blocks, staged residual entries)
// in device memory (so a recorded graph replays any batch), unitStart[j] the first unit of job j, a unit
// being WAVES * EPW elements. Persistent: workgroups stride over the units.
// Two wavefronts per SIMD are requested explicitly: the persistent loop's state would otherwise push the
// FP64 variants past 256 registers (one wavefront per SIMD); FP64 batch 56.6 ms against 59.3 ms with a few
// registers spilled.
template <int P, typename T, int WAVES, bool RES>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(2))) void element_tiled_batch(const SplitJob *__restrict__ jobs,
                                                                   const int *__restrict__ unitStart,
                                                                   const int *__restrict__ counts,
                                                                   const Basis1D basisIn) {
  using O = Ord<P>;
  __shared__ WaveLDSTile<P, T> wl[WAVES];
  __shared__ Basis1D basis;
  __shared__ double2 gref[O::Q * O::N];
  copy_basis(basis, basisIn);
  fill_gref<P>(gref, basisIn);
  __syncthreads();
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int numJobs = counts[0], units = counts[1];
  for (int u = blockIdx.x; u < units; u += gridDim.x) {
    const int j = __builtin_amdgcn_readfirstlane(find_job(unitStart, numJobs, u));  // uniform: scalar loads
    const SplitJob job = jobs[j];
    const int64_t first = (int64_t(u - unitStart[j]) * WAVES + wave) * O::EPW;
    tiled_wave<P, T, RES>(wl[wave], basis, gref, job, first, lane);
  }
}

// ---- NONLINEAR (Green) strain: element-parallel drop-in paths ------------------------------------------
// The reference's kernels with GreenStrainType::NONLINEAR (element/FEKernels.h:286-315, 585-1006): Green
// strain E = (Ux + Ux^T + Ux^T Ux) / 2, S = 2 mu E + lambda tr(E) I, weak residual s F S with F = I + Ux
// (the reverse sweep of the energy tr(E S) s / 2), and its tangent
//   K_ik(a, b) = sum_q s [mu (F F^T)_ik (n_a . n_b) + mu (F n_b)_i (F n_a)_k + lambda (F n_a)_i (F n_b)_k
//                         + delta_ik n_a^T S n_b],
// n_a the physical gradient of node a. FP64 only.
template <int P>
struct __align__(16) WaveLDSNL {
  using O = Ord<P>;
  static constexpr int NG = 20;  // doubles per quadrature point (see stage_geometry_nl)
  static constexpr int GEO_STRIDE = 2 * odd_units(O::Q * NG / 2, 16);
  union {
    double nodal[O::EPW][O::NODAL_STRIDE];  // x, y, u, v of each element node
    double2 pub[O::EPW][O::N][2];           // n_b and F n_b of the current quadrature point
  };
  double geo[O::EPW][GEO_STRIDE];
};

// Lane (el, q): geometry, deformation and stresses at quadrature point q, stored as
// [J^-1 (4), s, s F S (4), F (4), F F^T (xx, xy, yy), s S (xx, xy, yy), 0].
template <int P>
__device__ __forceinline__ void stage_geometry_nl(WaveLDSNL<P> &W, const Basis1D &B, const Material &m, int el,
                                                  int q) {
  constexpr int NP = Ord<P>::NP;
  const int qx = q % NP, qy = q / NP;
  double J00 = 0, J01 = 0, J10 = 0, J11 = 0, U00 = 0, U01 = 0, U10 = 0, U11 = 0;
#pragma unroll
  for (int ay = 0; ay < NP; ay++) {
#pragma unroll
    for (int ax = 0; ax < NP; ax++) {
      const int a = ay * NP + ax;
      const double gx = B.phi[ay][qy] * B.dphi[ax][qx];
      const double gy = B.phi[ax][qx] * B.dphi[ay][qy];
      const double2 xy = *reinterpret_cast<const double2 *>(&W.nodal[el][4 * a]);
      const double2 uv = *reinterpret_cast<const double2 *>(&W.nodal[el][4 * a + 2]);
      J00 = fma(xy.x, gx, J00);
      J01 = fma(xy.x, gy, J01);
      J10 = fma(xy.y, gx, J10);
      J11 = fma(xy.y, gy, J11);
      U00 = fma(uv.x, gx, U00);
      U01 = fma(uv.x, gy, U01);
      U10 = fma(uv.y, gx, U10);
      U11 = fma(uv.y, gy, U11);
    }
  }
  const double detJ = J00 * J11 - J01 * J10;
  const double inv = 1.0 / detJ;
  const double i00 = J11 * inv, i01 = -J01 * inv, i10 = -J10 * inv, i11 = J00 * inv;
  // displacement gradient Ux = dU/dxi J^-1
  const double G00 = U00 * i00 + U01 * i10, G01 = U00 * i01 + U01 * i11;
  const double G10 = U10 * i00 + U11 * i10, G11 = U10 * i01 + U11 * i11;
  const double exx = G00 + 0.5 * (G00 * G00 + G10 * G10);
  const double eyy = G11 + 0.5 * (G01 * G01 + G11 * G11);
  const double exy = 0.5 * (G01 + G10 + G00 * G01 + G10 * G11);
  const double tr = exx + eyy;
  const double Sxx = 2 * m.mu * exx + m.lambda * tr, Syy = 2 * m.mu * eyy + m.lambda * tr, Sxy = 2 * m.mu * exy;
  const double F00 = 1 + G00, F01 = G01, F10 = G10, F11 = 1 + G11;
  const double s = (B.w[qx] * B.w[qy]) * detJ * m.t;
  double *g = &W.geo[el][WaveLDSNL<P>::NG * q];
  g[0] = i00;
  g[1] = i01;
  g[2] = i10;
  g[3] = i11;
  g[4] = s;
  g[5] = s * (F00 * Sxx + F01 * Sxy);
  g[6] = s * (F00 * Sxy + F01 * Syy);
  g[7] = s * (F10 * Sxx + F11 * Sxy);
  g[8] = s * (F10 * Sxy + F11 * Syy);
  g[9] = F00;
  g[10] = F01;
  g[11] = F10;
  g[12] = F11;
  g[13] = F00 * F00 + F01 * F01;
  g[14] = F00 * F10 + F01 * F11;
  g[15] = F10 * F10 + F11 * F11;
  g[16] = s * Sxx;
  g[17] = s * Sxy;
  g[18] = s * Syy;
  g[19] = 0;
}

// Lane (el, b): residual of node b and (JAC) the blocks K(a, b) of every row a, in chunks of CH rows; the
// gradients of the chunk's rows come through the per-wave LDS one quadrature point at a time. All lanes of
// the wavefront must call this; `valid` masks the lane's writes. Every lane calls emit(at, k) for every row,
// in step, with a negative destination when it has none (so emit may exchange values across lanes).
template <int P, bool JAC, typename Where, typename Emit>
__device__ __forceinline__ void stage_stiffness_nl(WaveLDSNL<P> &W, const double2 *gref, const Material &m, int el,
                                                   int b, bool valid, double &Rx, double &Ry, Where &&where,
                                                   Emit &&emit) {
  using O = Ord<P>;
  constexpr int N = O::N, Q = O::Q, NG = WaveLDSNL<P>::NG, CH = 4;
  const int elc = el < O::EPW ? el : O::EPW - 1;
  const int bc = b < N ? b : 0;
  const double mu = m.mu, lam = m.lambda;
  Rx = 0;
  Ry = 0;
#pragma unroll 1
  for (int c0 = 0; c0 < (JAC ? N : 1); c0 += CH) {
    double A[CH][3], Tm[CH][4], H[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) {
      A[j][0] = A[j][1] = A[j][2] = 0;
      Tm[j][0] = Tm[j][1] = Tm[j][2] = Tm[j][3] = 0;
      H[j] = 0;
    }
#pragma unroll 1
    for (int q = 0; q < Q; q++) {
      const double *g = &W.geo[elc][NG * q];
      const double2 gb = gref[q * N + bc];
      const double nx = gb.x * g[0] + gb.y * g[2];
      const double ny = gb.x * g[1] + gb.y * g[3];
      if (c0 == 0) {
        Rx += g[5] * nx + g[6] * ny;
        Ry += g[7] * nx + g[8] * ny;
      }
      if (JAC) {
        const double s = g[4];
        const double px = g[9] * nx + g[10] * ny, py = g[11] * nx + g[12] * ny;
        const double spx = s * px, spy = s * py;
        const double qx = g[16] * nx + g[17] * ny, qy = g[17] * nx + g[18] * ny;
        if (valid) {
          W.pub[el][b][0] = make_double2(nx, ny);
          W.pub[el][b][1] = make_double2(px, py);
        }
        wave_sync();
#pragma unroll
        for (int j = 0; j < CH; j++) {
          const int a = c0 + j < N ? c0 + j : N - 1;
          const double2 na = W.pub[elc][a][0], pa = W.pub[elc][a][1];
          const double sd = s * (na.x * nx + na.y * ny);
          A[j][0] = fma(sd, g[13], A[j][0]);
          A[j][1] = fma(sd, g[14], A[j][1]);
          A[j][2] = fma(sd, g[15], A[j][2]);
          Tm[j][0] = fma(pa.x, spx, Tm[j][0]);
          Tm[j][1] = fma(pa.x, spy, Tm[j][1]);
          Tm[j][2] = fma(pa.y, spx, Tm[j][2]);
          Tm[j][3] = fma(pa.y, spy, Tm[j][3]);
          H[j] = fma(na.x, qx, fma(na.y, qy, H[j]));
        }
        wave_sync();  // the next point republishes into the same slots
      }
    }
    if (JAC) {  // every lane calls emit for every row, in step (negative destination: none)
      int at[CH];
#pragma unroll
      for (int j = 0; j < CH; j++) at[j] = valid && c0 + j < N ? where(c0 + j, b) : -1;
#pragma unroll
      for (int j = 0; j < CH; j++) {
        if (c0 + j >= N) break;
        // K_ik = mu A_ik + mu T_ki + lambda T_ik + delta_ik H, T_ik = sum_q s (F n_a)_i (F n_b)_k
        const double k[4] = {mu * A[j][0] + (mu + lam) * Tm[j][0] + H[j], mu * A[j][1] + mu * Tm[j][2] + lam * Tm[j][1],
                             mu * A[j][1] + mu * Tm[j][1] + lam * Tm[j][2], mu * A[j][2] + (mu + lam) * Tm[j][3] + H[j]};
        emit(at[j], k);
      }
    }
  }
}

template <int P, int WAVES, bool RES, bool JAC>
__global__ __launch_bounds__(WAVES * 64) void element_kernel_nl(const ElemArgs args) {
  using O = Ord<P>;
  constexpr int N = O::N;
  __shared__ WaveLDSNL<P> wl[WAVES];
  __shared__ Basis1D basis;
  __shared__ double2 gref[O::Q * O::N];
  __shared__ double stK[JAC ? WAVES : 1][4][64];  // Jacobian emission staging, as in element_kernel
  __shared__ int stAt[JAC ? WAVES : 1][64];
  copy_basis(basis, args.basis);
  fill_gref<P>(gref, args.basis);
  __syncthreads();

  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int el = lane / N, idx = lane % N;
  WaveLDSNL<P> &W = wl[wave];
  const int64_t slot = (int64_t(xcd_block()) * WAVES + wave) * O::EPW + el;
  const bool valid = lane < O::LANES && slot < args.count;
  int eid = 0, node = 0;
  if (valid) {
    eid = args.first + int(slot);
    node = args.conn[args.connPtr[eid] + idx];
    store_nodal(W, el, idx, *reinterpret_cast<const double2 *>(args.coords + 2 * int64_t(node)),
                *reinterpret_cast<const double2 *>(args.states + 2 * int64_t(node)));
  }
  wave_sync();
  if (valid) stage_geometry_nl<P>(W, basis, args.mat, el, idx);
  wave_sync();
  const int *mel = args.map + int64_t(eid) * N * N;
  double Rx, Ry;
  auto where = [&](int row, int col) __attribute__((always_inline)) { return mel[row * N + col]; };
  // Each add instruction covers the four components of 16 blocks (lanes 4i..4i+3 take block i).
  auto emit = [&](int at, const double(&k)[4]) __attribute__((always_inline)) {
    stAt[wave][lane] = at;
#pragma unroll
    for (int c = 0; c < 4; c++) stK[wave][c][lane] = k[c];
    wave_sync();
    const int c = lane & 3;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int src = (lane >> 2) + 16 * r;
      const int a = stAt[wave][src];
      if (a >= 0) unsafeAtomicAdd(args.values + a + c, stK[wave][c][src]);
    }
    wave_sync();
  };
  stage_stiffness_nl<P, JAC>(W, gref, args.mat, el, idx, valid, Rx, Ry, where, emit);
  if (RES && valid) {
    double *r = args.residual + 2 * int64_t(node);
    unsafeAtomicAdd(r, Rx);
    unsafeAtomicAdd(r + 1, Ry);
  }
}

// ---- owner-computes kernels (resident plans, write-once) -------------------------------------------
// One assembly of one resident plan: the plan's device arrays and the request's inputs/outputs.
struct MeshJob {
  const int *visitConn;   // [visits*N] nodes of each element visit, parts' visits contiguous
  const int *visitEid;    // [visits] element id of each visit
  const int *nodeLrow;    // [nodes] index of the node within its part's row list
  const int *rows;        // [nodes] owned rows, part by part, ascending within a part
  const int *rowLbase;    // [nodes] LDS block offset of each owned row within its part
  const int4 *runs;       // [runs] (LDS block start, global block start, blocks, 0)
  const int4 *partDesc;   // [parts][2] (first row, rows, blocks, first run), (runs, first visit, visits, 0)
  const uint32_t *rank;   // [E][N][RPAD] bytes: rank[e][b][a] = slot of column conn[e,b] within row conn[e,a]
  const int *colourOff;   // [parts][C+1] colour offsets within a part's visits (deterministic plans)
  const double *states;
  const double *coords;
  double *residual;       // may be NULL
  double *values;         // may be NULL
  double mu, lambda, t;
  int numParts, numColours;
  int accumulate;  // 1: the write-out adds to the outputs (+=) instead of storing them
  int pad_;
};

struct OwnerArgs {
  const MeshJob *jobs;    // device job table
  const int *jobStart;    // [numJobs+1] first global part of each job
  const int *counts;      // if non-NULL: device (numJobs, totalParts), read instead of the two fields below
  int numJobs;
  int totalParts;
  Basis1D basis;
};

struct PartDesc {
  int r0, nr, nb, rb, nruns, e0, ne;
};

__device__ __forceinline__ PartDesc part_desc(const MeshJob &J, int p) {
  const int4 a = cmem(J.partDesc)[2 * p], b = cmem(J.partDesc)[2 * p + 1];
  return PartDesc{a.x, a.y, a.z, a.w, b.x, b.y, b.z};
}

__device__ __forceinline__ Material job_material(const MeshJob &J) { return Material{J.mu, J.lambda, J.t}; }

// Per-lane gathered inputs of one element visit.
template <int P>
struct Gather {
  int node;
  int eid;
  int lrow;
  double2 xy, uv;
  uint32_t rk[Ord<P>::RPAD / 4];
};

// Level 1: node id and element id of visit `visit` (always in bounds: callers clamp the visit).
template <int P>
__device__ __forceinline__ void gather_level1(const MeshJob &J, int visit, int idx, Gather<P> &g) {
  g.node = load_once(gmem(J.visitConn) + int64_t(visit) * Ord<P>::N + idx);
  g.eid = load_once(gmem(J.visitEid) + visit);
}

// Level 2: coordinates, states, owner row and column ranks of the node/element of level 1.
template <int P>
__device__ __forceinline__ void gather_level2(const MeshJob &J, int idx, Gather<P> &g) {
  g.xy = gmem(reinterpret_cast<const double2 *>(J.coords))[g.node];
  g.uv = gmem(reinterpret_cast<const double2 *>(J.states))[g.node];
  g.lrow = gmem(J.nodeLrow)[g.node];
  constexpr int W = Ord<P>::RPAD / 4;
  const gptr<uint32_t> r = gmem(J.rank) + (int64_t(g.eid) * Ord<P>::N + idx) * W;
#pragma unroll
  for (int i = 0; i < W; i++) g.rk[i] = load_once(r + i);
}

// One chunk of the row-restricted stiffness: rows W.orow[el][base .. base+CH) (clamped to the element's
// count; rows past it are formed but not emitted). With `first`, the residual of node b is accumulated.
// Software-pipelined over quadrature points like stage_stiffness: iteration q publishes point q+1's
// weighted gradient of node b and issues the reads of the chunk's rows, then accumulates point q from
// values read one iteration earlier.
template <int P, typename T, bool JAC, int CH, bool ALL_LANES, typename Where, typename Emit>
__device__ __forceinline__ void rows_chunk(WaveLDS<P, T> &W, const double2 *gref, const Material &m, int el, int b,
                                           bool valid, int count, int base, bool first, double &Rx, double &Ry,
                                           Where &&where, Emit &&emit) {
  using O = Ord<P>;
  constexpr int Q = O::Q;
  using V2 = typename Vec2<T>::type;
  const T mu = T(m.mu), lam = T(m.lambda), c1 = lam + 2 * mu;
  const int elc = el < O::EPW ? el : O::EPW - 1;  // lanes past the last element read a real row
  const int bc = b < O::N ? b : 0;
  T S[CH][4];
  int arow[CH];
#pragma unroll
  for (int j = 0; j < CH; j++) {
    S[j][0] = S[j][1] = S[j][2] = S[j][3] = 0;
    arow[j] = W.orow[elc][min(base + j, max(count - 1, 0))];
  }
  // Point q's terms: residual (kept for the first chunk only) and the published weighted gradient.
  double rx = 0, ry = 0;
  auto accumulate = [&](const V2(&wc)[CH], T nx, T ny) __attribute__((always_inline)) {
    if (JAC) {
#pragma unroll
      for (int j = 0; j < CH; j++) {
        S[j][0] = fma(wc[j].x, nx, S[j][0]);
        S[j][1] = fma(wc[j].y, ny, S[j][1]);
        S[j][2] = fma(wc[j].x, ny, S[j][2]);
        S[j][3] = fma(wc[j].y, nx, S[j][3]);
      }
    }
  };
  auto step = [&](int q, const V2(&wc)[CH], V2(&wn)[CH], PointGeo &G, T &nx, T &ny) __attribute__((always_inline)) {
    T nx1, ny1;
    V2 v1;
    point_terms<T, V2>(G, nx1, ny1, v1, rx, ry);
    load_point<P, T>(W, gref, elc, bc, q + 2 < Q ? q + 2 : Q - 1, G);
    if (JAC) {
      const int buf = (q + 1) & 1;
      if (valid) W.sg[buf][el][b] = v1;
      wave_sync();
#pragma unroll
      for (int j = 0; j < CH; j++) wn[j] = W.sg[buf][elc][arow[j]];
    }
    accumulate(wc, nx, ny);
    nx = nx1;
    ny = ny1;
  };
  PointGeo G;
  load_point<P, T>(W, gref, elc, bc, 0, G);
  T nx, ny;
  V2 v;
  point_terms<T, V2>(G, nx, ny, v, rx, ry);
  load_point<P, T>(W, gref, elc, bc, Q > 1 ? 1 : 0, G);
  V2 wA[CH], wB[CH];
  if (JAC) {
    if (valid) W.sg[0][el][b] = v;
    wave_sync();
#pragma unroll
    for (int j = 0; j < CH; j++) wA[j] = W.sg[0][elc][arow[j]];
  }
  constexpr int PAIRS = (Q - 1) / 2;
#pragma unroll 1
  for (int i = 0; i < PAIRS; i++) {
    step(2 * i, wA, wB, G, nx, ny);
    step(2 * i + 1, wB, wA, G, nx, ny);
  }
  if constexpr ((Q - 1) % 2 == 1) {
    step(Q - 2, wA, wB, G, nx, ny);
    accumulate(wB, nx, ny);
  } else {
    accumulate(wA, nx, ny);
  }
  if (first) {
    Rx += rx;
    Ry += ry;
  }
  if (JAC && (ALL_LANES || valid)) {
    int at[CH];
#pragma unroll
    for (int j = 0; j < CH; j++) at[j] = valid && base + j < count ? where(arow[j], b) : -1;
#pragma unroll
    for (int j = 0; j < CH; j++) {
      if (base + j >= count) break;
      double k[4];
      k[0] = double(c1 * S[j][0] + mu * S[j][1]);
      k[1] = double(lam * S[j][2] + mu * S[j][3]);
      k[2] = double(lam * S[j][3] + mu * S[j][2]);
      k[3] = double(c1 * S[j][1] + mu * S[j][0]);
      emit(at[j], k);
    }
  }
  wave_sync();  // the next chunk republishes into the same buffers
}

// Row-restricted stiffness (orders >= 3, where an element is visited by several parts): lane (el, b) forms
// K(a, b) only for the rows a this part owns, W.orow[el][0..count), in chunks of 8 rows (4 in FP64; maxCount,
// the largest count over the wavefront's elements, sets the number of chunks). Over all parts visiting an
// element this forms each row of K_e once, instead of the whole symmetric K_e per visit. The residual of
// node b is accumulated in the first chunk. where and emit as for stage_stiffness, except that emit gets a
// negative destination for rows it must skip. With ALL_LANES every lane calls emit for every row of a
// chunk, in step (count must then be the same for all lanes), so emit may exchange values across lanes.
template <int P, typename T, bool JAC, bool ALL_LANES, typename Where, typename Emit>
__device__ __forceinline__ void stage_stiffness_rows(WaveLDS<P, T> &W, const double2 *gref, const Material &m, int el,
                                                     int b, bool valid, int count, int maxCount, double &Rx,
                                                     double &Ry, Where &&where, Emit &&emit) {
  constexpr int CH = sizeof(T) == 8 ? 4 : 8;  // FP64 chunks are halved to stay within 256 registers
  Rx = 0;
  Ry = 0;
  const int chunks = JAC ? uniform(max((maxCount + CH - 1) / CH, 1)) : 1;
#pragma unroll 1
  for (int c = 0; c < chunks; c++) {
    rows_chunk<P, T, JAC, CH, ALL_LANES>(W, gref, m, el, b, valid, count, c * CH, c == 0, Rx, Ry, where, emit);
  }
}

typedef float floatx4 __attribute__((ext_vector_type(4)));

// MFMA form of the row-restricted stiffness (mixed precision, orders >= 3). For each chunk of 8 owned rows
// of an element, the Gram sums are one product C = A B with
//   A[(a, alpha)][q] = s_q dN_a/dx_alpha(q)   (16 rows: 8 nodes x 2 components),
//   B[q][(b, beta)]  = dN_b/dx_beta(q)        (2N columns, in tiles of 16),
// so C[(a, alpha)][(b, beta)] = S_alpha,beta(a, b), accumulated over the quadrature points four at a time
// by v_mfma_f32_16x16x4_f32 (lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds
// C[4 (l / 16) + r][l % 16], r = 0..3). Each step, lane (el, b) writes dN_b/dx at four points into LDS; a
// wavefront then runs the products of each of its elements in turn. Lane pairs (l, l ^ 1) hold the x and
// y columns of one node b, so after one swap each lane has the four sums of one block (a, b) and emits it.
// The residual of node b is accumulated in FP64 during the first chunk.
template <int P, int MAXB, bool DET, typename WL>
__device__ __forceinline__ void stage_stiffness_mfma(WL &W, const double2 *gref, const Material &m, int el, int b,
                                                     unsigned long long ownMask, double &Rx, double &Ry, double *acc) {
  using O = Ord<P>;
  constexpr int N = O::N, Q = O::Q, EPW = O::EPW;
  constexpr int KSTEPS = (Q + 3) / 4, TILES = (2 * N + 15) / 16;
  constexpr unsigned long long elMask = (1ull << N) - 1;
  const int lane = int(threadIdx.x % 64);
  const bool colLane = el < EPW;  // lane (el, b) holds a column node of element el
  const float mu = float(m.mu), lam = float(m.lambda), c1 = lam + 2 * mu;
  int cnt[EPW];
  int maxCount = 0;
#pragma unroll
  for (int e = 0; e < EPW; e++) {
    cnt[e] = __popcll((ownMask >> (e * N)) & elMask);
    maxCount = max(maxCount, cnt[e]);
  }
  Rx = 0;
  Ry = 0;
  const int iA = lane % 16, kk = lane / 16;
  const int chunks = max((maxCount + 7) / 8, 1);
#pragma unroll 1
  for (int c = 0; c < chunks; c++) {
    floatx4 D[EPW][TILES];
    int arowA[EPW];
#pragma unroll
    for (int e = 0; e < EPW; e++) {
#pragma unroll
      for (int t = 0; t < TILES; t++) D[e][t] = floatx4{0.f, 0.f, 0.f, 0.f};
      arowA[e] = W.orow[e][min(c * 8 + iA / 2, max(cnt[e] - 1, 0))];
    }
    // Step st: lane (el, b) writes dN_b/dx at points 4st .. 4st+3 (zero past Q); the residual is added in
    // the first chunk.
    auto write_slice = [&](int st) __attribute__((always_inline)) {
      if (colLane) {
#pragma unroll
        for (int qq = 0; qq < 4; qq++) {
          const int q = 4 * st + qq;
          float2 n = make_float2(0.f, 0.f);
          if (q < Q) {
            const double *g = &W.geo[el][8 * q];
            const double2 gg = gref[q * N + b];
            const double nxd = gg.x * g[0] + gg.y * g[2];
            const double nyd = gg.x * g[1] + gg.y * g[3];
            if (c == 0) {
              Rx += g[5] * nxd + g[7] * nyd;
              Ry += g[7] * nxd + g[6] * nyd;
            }
            n = make_float2(float(nxd), float(nyd));
          }
          W.ns[el][qq][b] = n;
        }
      }
    };
    // MFMA operands of step st for every element of the wavefront.
    struct Operands {
      float a[EPW];
      float bv[EPW][TILES];
    };
    auto read_operands = [&](int st, Operands &op) __attribute__((always_inline)) {
      const int q = 4 * st + kk;
#pragma unroll
      for (int e = 0; e < EPW; e++) {
        op.a[e] = 0.f;
        if (q < Q) {
          const float2 na = W.ns[e][kk][arowA[e]];
          op.a[e] = float(W.geo[e][8 * q + 4]) * ((iA & 1) ? na.y : na.x);
        }
#pragma unroll
        for (int t = 0; t < TILES; t++) {
          const int col = 16 * t + iA;
          op.bv[e][t] = 0.f;
          if (col < 2 * N) {
            const float2 nb = W.ns[e][kk][col / 2];
            op.bv[e][t] = (col & 1) ? nb.y : nb.x;
          }
        }
      }
    };
    auto products = [&](const Operands &op) __attribute__((always_inline)) {
#pragma unroll
      for (int e = 0; e < EPW; e++) {
#pragma unroll
        for (int t = 0; t < TILES; t++) {
          D[e][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(op.a[e], op.bv[e][t], D[e][t], 0, 0, 0);
        }
      }
    };
    // Pipeline: the slice of step st+1 is written and read into registers while step st's products run;
    // one LDS slice suffices because a wavefront's LDS accesses execute in order.
    Operands opA, opB;
    write_slice(0);
    wave_sync();
    read_operands(0, opA);
    auto step = [&](int st, const Operands &cur, Operands &nxt) __attribute__((always_inline)) {
      wave_sync();  // cur was read before the slice is rewritten (program order)
      write_slice(st + 1);
      wave_sync();
      read_operands(st + 1, nxt);
      products(cur);
    };
    constexpr int PAIRS = (KSTEPS - 1) / 2;
#pragma unroll 1
    for (int i = 0; i < PAIRS; i++) {
      step(2 * i, opA, opB);
      step(2 * i + 1, opB, opA);
    }
    if constexpr ((KSTEPS - 1) % 2 == 1) {
      step(KSTEPS - 2, opA, opB);
      products(opB);
    } else {
      products(opA);
    }
    wave_sync();  // the next chunk rewrites the slice
    const bool even = (lane & 1) == 0;
    const int jr = 2 * (lane / 16) + (even ? 0 : 1);  // chunk row this lane emits
    // Destinations of this lane's blocks, all looked up before the first add.
    int at[EPW][TILES];
#pragma unroll
    for (int e = 0; e < EPW; e++) {
      const bool rowOk = c * 8 + jr < cnt[e];
      const int arow = W.orow[e][rowOk ? c * 8 + jr : 0];
#pragma unroll
      for (int t = 0; t < TILES; t++) {
        const int bcol = (16 * t + lane % 16) / 2;
        const int at0 = W.obase[e][arow] + rank_of(W, e, arow, bcol < N ? bcol : 0);
        at[e][t] = rowOk && bcol < N ? at0 : -1;
      }
    }
#pragma unroll
    for (int e = 0; e < EPW; e++) {
#pragma unroll
      for (int t = 0; t < TILES; t++) {
        const floatx4 d = D[e][t];
        float o[4];
#pragma unroll
        for (int r = 0; r < 4; r++) o[r] = __shfl_xor(d[r], 1);
        const float sxx = even ? d[0] : o[2], syx = even ? d[1] : o[3];
        const float sxy = even ? o[0] : d[2], syy = even ? o[1] : d[3];
        if (at[e][t] >= 0) {
          double *dst = acc + at[e][t];
          const double k[4] = {double(c1 * sxx + mu * syy), double(lam * sxy + mu * syx),
                               double(lam * syx + mu * sxy), double(c1 * syy + mu * sxx)};
          if (DET) {
#pragma unroll
            for (int cc = 0; cc < 4; cc++) dst[cc * acc_plane<MAXB>()] += k[cc];
          } else {
#pragma unroll
            for (int cc = 0; cc < 4; cc++) atomicAdd(dst + cc * acc_plane<MAXB>(), k[cc]);
          }
        }
      }
    }
  }
}

// The accumulator holds a part's BSR blocks as four planes, acc[c * acc_plane + block] for block component c
// (row-major xx, xy, yx, yy): the lanes of one LDS atomic then spread over all banks.
//
// An element batch runs in three phases: batch_nodes stages the gathered nodes and column ranks,
// batch_geometry forms each quadrature point's geometry (both use only the wavefront's own LDS), and
// batch_stiffness marks the part's rows, forms the residual and stiffness and adds them to the part's
// accumulators. The streaming kernel runs the previous part's write-out between the first two phases.
template <int P, typename T>
__device__ __forceinline__ void batch_nodes(WaveLDS<P, T> &W, const Gather<P> &g, bool valid, int el, int idx) {
  if (valid) {
    store_nodal(W, el, idx, g.xy, g.uv);
#pragma unroll
    for (int i = 0; i < Ord<P>::RPAD / 4; i++) W.rk[el][idx][i] = g.rk[i];
  }
  wave_sync();
}

template <int P, typename T>
__device__ __forceinline__ void batch_geometry(WaveLDS<P, T> &W, const Basis1D &basis, const Material &mat, bool valid,
                                               int el, int idx) {
  if (valid) stage_geometry<P, T>(W, basis, mat, el, idx);
  wave_sync();
}

template <int P, typename T, int MAXB, bool RES, bool JAC, bool DET>
__device__ __forceinline__ void batch_stiffness(const Material &mat, WaveLDS<P, T> &W, const double2 *gref,
                                                double *acc, double *racc, const int *lrows, const int *lbase, int nr,
                                                const Gather<P> &g, bool valid, int el, int idx) {
  using O = Ord<P>;
  int own = -1;
  if (valid) {
    own = (g.lrow < nr && lrows[g.lrow] == g.node) ? g.lrow : -1;
    W.obase[el][idx] = own >= 0 ? lbase[own] : -1;
  }
  // Orders >= 3: compact this element's owned rows (lanes of element e are e*N .. e*N+N-1).
  int count = 0, maxCount = 0;
  unsigned long long ownMask = 0;
  if constexpr (P >= 3) {
    const unsigned long long m = __ballot(valid && own >= 0);
    ownMask = m;
    constexpr unsigned long long elMask = (1ull << O::N) - 1;
#pragma unroll
    for (int e = 0; e < O::EPW; e++) maxCount = max(maxCount, __popcll((m >> (e * O::N)) & elMask));
    const unsigned long long mine = el < O::EPW ? (m >> (el * O::N)) & elMask : 0ull;
    count = __popcll(mine);
    if (valid && own >= 0) W.orow[el][__popcll(mine & ((1ull << idx) - 1))] = (unsigned char)idx;
  }
  wave_sync();
  double Rx, Ry;
  auto where = [&](int row, int col) __attribute__((always_inline)) {
    const int ob = W.obase[el][row];
    const int rk = rank_of(W, el, row, col);  // read unconditionally: no branch between the lookups
    return ob >= 0 ? ob + rk : -1;
  };
  auto emit = [&](int at, const double(&k)[4]) __attribute__((always_inline)) {
    if (at < 0) return;
    double *dst = acc + at;
    if (DET) {
#pragma unroll
      for (int c = 0; c < 4; c++) dst[c * acc_plane<MAXB>()] += k[c];
    } else {
#pragma unroll
      for (int c = 0; c < 4; c++) atomicAdd(dst + c * acc_plane<MAXB>(), k[c]);
    }
  };
  if constexpr (P >= 3 && JAC && std::is_same<T, float>::value) {
    stage_stiffness_mfma<P, MAXB, DET>(W, gref, mat, el, idx, ownMask, Rx, Ry, acc);
  } else if constexpr (P >= 3) {
    stage_stiffness_rows<P, T, JAC, false>(W, gref, mat, el, idx, valid, count, maxCount, Rx, Ry, where, emit);
  } else {
    stage_stiffness<P, T, JAC>(W, gref, mat, el, idx, valid, Rx, Ry, where, emit);
  }
  if (RES && valid && own >= 0) {
    if (DET) {
      racc[2 * own] += double(Rx);
      racc[2 * own + 1] += double(Ry);
    } else {
      atomicAdd(&racc[2 * own], double(Rx));
      atomicAdd(&racc[2 * own + 1], double(Ry));
    }
  }
  wave_sync();
}

template <int P, typename T, int MAXB, bool RES, bool JAC, bool DET>
__device__ __forceinline__ void owner_batch(const Material &mat, WaveLDS<P, T> &W, const Basis1D &basis,
                                            const double2 *gref, double *acc, double *racc, const int *lrows,
                                            const int *lbase, int nr, const Gather<P> &g, bool valid, int el,
                                            int idx) {
  batch_nodes<P, T>(W, g, valid, el, idx);
  batch_geometry<P, T>(W, basis, mat, valid, el, idx);
  batch_stiffness<P, T, MAXB, RES, JAC, DET>(mat, W, gref, acc, racc, lrows, lbase, nr, g, valid, el, idx);
}

// A part's write run: LDS block start, global block start, blocks. In LDS either as int4 (x, y, z) or
// packed as int2 (x = LDS start | blocks << 16, y = global start).
struct Run {
  int lds, global, len;
};
__device__ __forceinline__ Run unpack_run(const int4 &r) { return Run{r.x, r.y, r.z}; }
__device__ __forceinline__ Run unpack_run(const int2 &r) { return Run{r.x & 0xffff, r.y, r.x >> 16}; }

// Writes a part's accumulated BSR runs (one wave per run) and owned residual rows. With HALVES, lane pairs
// move the two 16-byte halves of a block (contiguous store instructions; faster at orders 1-2), else each
// lane moves a whole block (faster at orders 3-4, whose write-out loops are shorter). With ZERO, every
// value read is cleared, so the accumulators are all zero again afterwards (blocks past the part's count
// were never touched).
// Accumulate-mode write-out (J.accumulate: the reference kernels' += semantics). Lane pairs move block halves
// as in the HALVES form; each lane takes up to eight store slots per round and loads all their old values
// before adding and storing any, so a round's loads are in flight together.
template <bool RES, bool JAC, int WAVES, int MAXB, bool ZERO, typename RunT>
__device__ void owner_write_accumulate(const MeshJob &J, double *acc, double *racc, const RunT *lruns, int nruns,
                                       const int *lrows, int nr) {
  constexpr int U = 8;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (JAC && J.values) {
    const gptr_mut<double2> values = gmem_mut(reinterpret_cast<double2 *>(J.values));
    for (int r = wave; r < nruns; r += WAVES) {
      const Run run = unpack_run(lruns[r]);
      const gptr_mut<double2> dst = values + 2 * int64_t(run.global);
      for (int base = lane; base < 2 * run.len; base += 64 * U) {
        double2 old[U];
        double v0[U], v1[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
          const int i = base + 64 * k;
          if (i < 2 * run.len) {
            old[k] = dst[i];
            const double *src = acc + 2 * (i & 1) * acc_plane<MAXB>() + run.lds + (i >> 1);
            v0[k] = src[0];
            v1[k] = src[acc_plane<MAXB>()];
          }
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
          const int i = base + 64 * k;
          if (i < 2 * run.len) {
            store_once(dst + i, v0[k] + old[k].x, v1[k] + old[k].y);
            if (ZERO) {
              double *src = acc + 2 * (i & 1) * acc_plane<MAXB>() + run.lds + (i >> 1);
              src[0] = 0;
              src[acc_plane<MAXB>()] = 0;
            }
          }
        }
      }
    }
  }
  if (RES && J.residual) {
    const gptr_mut<double2> res = gmem_mut(reinterpret_cast<double2 *>(J.residual));
    for (int i = int(threadIdx.x); i < nr; i += WAVES * 64) {
      const double2 a = reinterpret_cast<const double2 *>(racc)[i], old = res[lrows[i]];
      res[lrows[i]] = make_double2(a.x + old.x, a.y + old.y);
      if (ZERO) reinterpret_cast<double2 *>(racc)[i] = make_double2(0, 0);
    }
  }
}

template <bool RES, bool JAC, int WAVES, int MAXB, bool ZERO = false, bool HALVES = false, typename RunT>
__device__ __forceinline__ void owner_write(const MeshJob &J, double *acc, double *racc, const RunT *lruns, int nruns,
                                            const int *lrows, int nr) {
  if (J.accumulate) {
    owner_write_accumulate<RES, JAC, WAVES, MAXB, ZERO>(J, acc, racc, lruns, nruns, lrows, nr);
    return;
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (JAC && J.values) {
    const gptr_mut<double2> values = gmem_mut(reinterpret_cast<double2 *>(J.values));
    for (int r = wave; r < nruns; r += WAVES) {
      const Run run = unpack_run(lruns[r]);
      const gptr_mut<double2> dst = values + 2 * int64_t(run.global);
      if (HALVES) {
        // Lane pair (2k, 2k+1) moves the two 16-byte halves of block k, so each store instruction writes
        // 1 KB contiguously.
        for (int i = lane; i < 2 * run.len; i += 64) {
          double *src = acc + 2 * (i & 1) * acc_plane<MAXB>() + run.lds + (i >> 1);
          store_once(dst + i, src[0], src[acc_plane<MAXB>()]);
          if (ZERO) {
            src[0] = 0;
            src[acc_plane<MAXB>()] = 0;
          }
        }
      } else {
        // One block per lane (two 16-byte stores).
        double *src = acc + run.lds;
        for (int i = lane; i < run.len; i += 64) {
          store_once(dst + 2 * i, src[i], src[i + acc_plane<MAXB>()]);
          store_once(dst + 2 * i + 1, src[i + 2 * acc_plane<MAXB>()], src[i + 3 * acc_plane<MAXB>()]);
          if (ZERO) {
#pragma unroll
            for (int c = 0; c < 4; c++) src[i + c * acc_plane<MAXB>()] = 0;
          }
        }
      }
    }
  }
  if (RES && J.residual) {
    const gptr_mut<double2> res = gmem_mut(reinterpret_cast<double2 *>(J.residual));
    for (int i = int(threadIdx.x); i < nr; i += WAVES * 64) {
      res[lrows[i]] = reinterpret_cast<const double2 *>(racc)[i];
      if (ZERO) reinterpret_cast<double2 *>(racc)[i] = make_double2(0, 0);
    }
  }
}

// Zeroes the first nb blocks of each accumulator plane.
template <int THREADS, int MAXB>
__device__ __forceinline__ void zero_planes(double *acc, int nb) {
  const int h = (nb + 1) / 2;  // MAXB is even, so the rounded-up pair stays inside the plane
#pragma unroll
  for (int c = 0; c < 4; c++) {
    double2 *plane = reinterpret_cast<double2 *>(acc + c * acc_plane<MAXB>());
    for (int i = int(threadIdx.x); i < h; i += THREADS) plane[i] = make_double2(0, 0);
  }
}

// Per-part kernel of a single job (deterministic plans and parts that take several batches per
// wavefront): one workgroup per part.
template <int P, typename T, int WAVES, int MAXB, int MAXR, bool RES, bool JAC, bool DET>
__global__ __launch_bounds__(WAVES * 64) void owner_kernel(const OwnerArgs args) {
  using O = Ord<P>;
  __shared__ WaveLDS<P, T> wl[WAVES];
  __shared__ Basis1D basis;
  __shared__ double2 gref[O::Q * O::N];
  __shared__ __align__(16) double acc[JAC ? 4 * acc_plane<MAXB>() : 2];
  __shared__ __align__(16) double racc[RES ? 2 * MAXR : 2];
  __shared__ int lrows[MAXR];
  __shared__ int lbase[MAXR];
  __shared__ int4 lruns[JAC ? MAXR : 1];

  const MeshJob &J = args.jobs[0];
  // XCD-aware order: consecutive workgroup ids land on different XCDs, so give each XCD a
  // contiguous range of (spatially adjacent) parts to keep shared element data in its L2.
  const int per = int(gridDim.x) / 8;
  const int part = int(blockIdx.x % 8) * per + int(blockIdx.x / 8);
  if (part >= args.totalParts) return;

  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int el = lane / O::N, idx = lane % O::N;
  const bool laneOk = lane < O::LANES;
  const PartDesc d = part_desc(J, part);
  const int vlast = d.e0 + d.ne - 1;  // every part has at least one visit
  constexpr int STEP = WAVES * O::EPW;
  const Material mat = job_material(J);

  for (int i = threadIdx.x; i < d.nr; i += blockDim.x) {
    lrows[i] = J.rows[d.r0 + i];
    lbase[i] = J.rowLbase[d.r0 + i];
  }
  if (JAC) {
    for (int i = threadIdx.x; i < d.nruns; i += blockDim.x) lruns[i] = J.runs[d.rb + i];
    zero_planes<WAVES * 64, MAXB>(acc, d.nb);
  }
  if (RES) {
    for (int i = threadIdx.x; i < d.nr; i += blockDim.x) reinterpret_cast<double2 *>(racc)[i] = make_double2(0, 0);
  }
  copy_basis(basis, args.basis);
  fill_gref<P>(gref, args.basis);
  __syncthreads();

  WaveLDS<P, T> &W = wl[wave];
  auto batch = [&](const Gather<P> &g, bool ok) __attribute__((always_inline)) {
    owner_batch<P, T, MAXB, RES, JAC, DET>(mat, W, basis, gref, acc, racc, lrows, lbase, d.nr, g, ok, el, idx);
  };
  if (DET) {
    const int *off = J.colourOff + int64_t(part) * (J.numColours + 1);
    for (int c = 0; c < J.numColours; c++) {
      const int cb = off[c], ce = off[c + 1];
      for (int b = cb + wave * O::EPW; b < ce; b += STEP) {
        const bool ok = laneOk && b + el < ce;
        Gather<P> g;
        gather_level1<P>(J, min(d.e0 + b + el, vlast), idx, g);
        gather_level2<P>(J, idx, g);
        batch(g, ok);
      }
      __syncthreads();
    }
  } else {
    // Software pipeline: gathers for batch k+1 (coordinates, states, ranks) and k+2 (node ids) are in
    // flight while batch k is computed. Loads are unconditional, indices clamped into the part.
    int vcur = wave * O::EPW + el;
    Gather<P> cur, nxt;
    gather_level1<P>(J, min(d.e0 + vcur, vlast), idx, cur);
    gather_level2<P>(J, idx, cur);
    gather_level1<P>(J, min(d.e0 + vcur + STEP, vlast), idx, nxt);
    for (; vcur - el < d.ne; vcur += STEP) {
      const bool ok = laneOk && vcur < d.ne;
      const Gather<P> g = cur;
      cur = nxt;
      gather_level2<P>(J, idx, cur);
      gather_level1<P>(J, min(d.e0 + vcur + 2 * STEP, vlast), idx, nxt);
      batch(g, ok);
    }
    __syncthreads();
  }
  owner_write<RES, JAC, WAVES, MAXB, false, (P <= 2)>(J, acc, racc, lruns, d.nruns, lrows, d.nr);
}

// One thread's entries of a part's row tables: owned row `tid` (node, LDS block base) and run `tid`.
struct PartTables {
  int row, lbase;
  int runLds, runGlobal, runLen;
};

__device__ __forceinline__ void load_part_tables(const MeshJob &J, const PartDesc &d, int tid, PartTables &t) {
  const int i = min(tid, d.nr - 1);
  t.row = gmem(J.rows)[d.r0 + i];
  t.lbase = gmem(J.rowLbase)[d.r0 + i];
  const int4 r = gmem(J.runs)[d.rb + min(tid, d.nruns - 1)];
  t.runLds = r.x;
  t.runGlobal = r.y;
  t.runLen = r.z;
}

// Parts [p0, p1) of one job, software-pipelined. The row and run tables are double-buffered by part parity,
// so a step needs two barriers. Step p stores part p's tables, stages its nodes (gathered during the
// previous step), writes out part p-1 (its stores drain while the geometry forms) and forms part p's
// geometry; after barrier A it issues part p+1's coordinate/state/rank gathers (their node ids arrived
// during the previous step), part p+2's node ids, part p+1's tables and part p+3's descriptor, and computes
// part p's stiffness from registers and LDS only; barrier C then closes part p's accumulators. The barriers
// wait for LDS traffic only, so no wait on HBM falls inside the compute phases.
template <int P, typename T, int WAVES, int MAXB, int MAXR, bool RES, bool JAC>
__device__ __forceinline__ void stream_parts(const MeshJob &J, int p0, int p1, WaveLDS<P, T> &W, const Basis1D &basis,
                                             const double2 *gref, double *acc, double *racc, int (*lrows)[MAXR],
                                             int *lbase, int2 (*lruns)[MAXR]) {
  using O = Ord<P>;
  constexpr int THREADS = WAVES * 64;
  const int tid = threadIdx.x, lane = tid % 64;
  const int el = lane / O::N, idx = lane % O::N;
  const bool laneOk = lane < O::LANES;
  const int myv = (tid / 64) * O::EPW + el;  // this lane's element visit within a part
  const Material mat = job_material(J);
  const int plast = p1 - 1;

  PartDesc dc = part_desc(J, p0);
  PartDesc dn = part_desc(J, min(p0 + 1, plast));
  PartDesc dnn = part_desc(J, min(p0 + 2, plast));
  PartTables tc;
  load_part_tables(J, dc, tid, tc);
  Gather<P> gc, gn;
  gather_level1<P>(J, dc.e0 + min(myv, dc.ne - 1), idx, gc);
  gather_level1<P>(J, dn.e0 + min(myv, dn.ne - 1), idx, gn);
  gather_level2<P>(J, idx, gc);

  // The accumulators start zeroed; each write-out clears what it reads, so they are zero at every part's
  // start. (The previous job's write-out finished before this call; the first barrier below orders it.)
  lds_barrier();
  if (JAC) zero_planes<THREADS, MAXB>(acc, MAXB);
  if (RES) {
    for (int i = tid; i < MAXR; i += THREADS) reinterpret_cast<double2 *>(racc)[i] = make_double2(0, 0);
  }
  PartDesc dprev = dc;
  for (int p = p0; p < p1; p++) {
    const int tb = (p - p0) & 1;  // table buffer of part p; part p-1's is tb ^ 1
    const bool ok = laneOk && myv < dc.ne;
    // Part p's tables: their buffer was last read by part p-2's write-out, and lbase by part p-1's
    // stiffness, both before the previous step's closing barrier.
    if (tid < dc.nr) {
      lrows[tb][tid] = tc.row;
      lbase[tid] = tc.lbase;
    }
    if (JAC && tid < dc.nruns) lruns[tb][tid] = make_int2(tc.runLds | (tc.runLen << 16), tc.runGlobal);
    // Stage this part's nodes, then write out the previous part: its stores drain while this part's
    // geometry forms (both phases use only the wavefront's own LDS besides the previous part's
    // accumulators and tables).
    batch_nodes<P, T>(W, gc, ok, el, idx);
    if (p > p0) {
      owner_write<RES, JAC, WAVES, MAXB, true, (P <= 2)>(J, acc, racc, lruns[tb ^ 1], dprev.nruns, lrows[tb ^ 1],
                                                         dprev.nr);
    }
    batch_geometry<P, T>(W, basis, mat, ok, el, idx);
    lds_barrier();  // A: part p's tables are stored; part p-1's accumulators are written out and cleared
    // Prefetch (all loads of this step are issued here, before the stiffness phase).
    const PartDesc d3 = part_desc(J, min(p + 3, plast));
    Gather<P> gnn;
    gather_level2<P>(J, idx, gn);
    PartTables tn;
    load_part_tables(J, dn, tid, tn);
    gather_level1<P>(J, dnn.e0 + min(myv, dnn.ne - 1), idx, gnn);

    batch_stiffness<P, T, MAXB, RES, JAC, false>(mat, W, gref, acc, racc, lrows[tb], lbase, dc.nr, gc, ok, el, idx);
    lds_barrier();  // C: the part's accumulators are complete

    dprev = dc;
    dc = dn;
    dn = dnn;
    dnn = d3;
    tc = tn;
    gc = gn;
    gn = gnn;
  }
  const int tl = (p1 - 1 - p0) & 1;
  owner_write<RES, JAC, WAVES, MAXB, false, (P <= 2)>(J, acc, racc, lruns[tl], dprev.nruns, lrows[tl], dprev.nr);
}

// Streaming kernel for parts that take one batch per wavefront (the default partition), over one or
// many jobs (a batch of resident meshes of one order). A persistent workgroup walks a contiguous range
// of global parts, job by job; job fields and material sit in scalar registers for the job's parts.
template <int P, typename T, int WAVES, int MAXB, int MAXR, bool RES, bool JAC>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(2, 2)))
void owner_stream_kernel(const OwnerArgs args) {
  using O = Ord<P>;
  static_assert(MAXR <= WAVES * 64, "one owned row per thread");
  static_assert(MAXB < 32768, "packed runs hold LDS starts and lengths in 16 bits");
  __shared__ WaveLDS<P, T> wl[WAVES];
  __shared__ Basis1D basis;
  __shared__ double2 gref[O::Q * O::N];
  __shared__ __align__(16) double acc[JAC ? 4 * acc_plane<MAXB>() : 2];
  __shared__ __align__(16) double racc[RES ? 2 * MAXR : 2];
  __shared__ int lrows[2][MAXR];
  __shared__ int lbase[MAXR];
  __shared__ int2 lruns[JAC ? 2 : 1][MAXR];

  // Contiguous range of global parts per workgroup; workgroup w runs on XCD w % 8, and each XCD's
  // workgroups share one contiguous (spatially adjacent, Morton-ordered) stretch of parts.
  const int xcd = int(blockIdx.x % 8), slot = int(blockIdx.x / 8), perX = int(gridDim.x / 8);
  const int numJobs = uniform(args.counts ? args.counts[0] : args.numJobs);
  const int64_t np = uniform(args.counts ? args.counts[1] : args.totalParts);
  const int xb = int(np * xcd / 8), xe = int(np * (xcd + 1) / 8);
  const int pb = xb + int(int64_t(xe - xb) * slot / perX);
  const int pe = xb + int(int64_t(xe - xb) * (slot + 1) / perX);
  if (pb >= pe) return;

  copy_basis(basis, args.basis);
  fill_gref<P>(gref, args.basis);
  WaveLDS<P, T> &W = wl[threadIdx.x / 64];

  // job of the first part: binary search of the job start table
  int j = 0;
  {
    int lo = 0, hi = numJobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (args.jobStart[mid] <= pb) lo = mid;
      else hi = mid - 1;
    }
    j = uniform(lo);
  }
  for (int gp = pb; gp < pe; j++) {
    const int js = uniform(args.jobStart[j]);
    const int je = uniform(min(args.jobStart[j + 1], pe));
    const MeshJob J = args.jobs[j];
    if (je > gp) {
      stream_parts<P, T, WAVES, MAXB, MAXR, RES, JAC>(J, gp - js, je - js, W, basis, gref, acc, racc, lrows, lbase,
                                                      lruns);
      gp = je;
    }
  }
}

// ---- BSR SpMV ---------------------------------------------------------------------------------------
// One wavefront per block row; lanes stride over the row's blocks and reduce.
__global__ __launch_bounds__(256) void bsr_spmv_kernel(int numRows, const int *rowp, const int *cols,
                                                       const double *values, const double *x, double *y) {
  const int r = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (r >= numRows) return;
  double y0 = 0, y1 = 0;
  for (int k = rowp[r] + lane; k < rowp[r + 1]; k += 64) {
    const double4 A = *reinterpret_cast<const double4 *>(values + 4 * int64_t(k));
    const double2 xv = *reinterpret_cast<const double2 *>(x + 2 * int64_t(cols[k]));
    y0 += A.x * xv.x + A.y * xv.y;
    y1 += A.z * xv.x + A.w * xv.y;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    y0 += __shfl_xor(y0, off);
    y1 += __shfl_xor(y1, off);
  }
  if (lane == 0) *reinterpret_cast<double2 *>(y + 2 * int64_t(r)) = make_double2(y0, y1);
}

// ---- host helpers -------------------------------------------------------------------------------------
namespace {

int make_basis(int order, Basis1D &b) {
  std::memset(&b, 0, sizeof(b));
  double phi[25], dphi[25], xg[5], wg[5];
  int rc = fem_basis_table(order, phi, dphi, xg, wg);
  if (rc) return rc;
  const int np = order + 1;
  for (int k = 0; k < np; k++) {
    for (int i = 0; i < np; i++) {
      b.phi[k][i] = phi[k * np + i];
      b.dphi[k][i] = dphi[k * np + i];
    }
  }
  for (int i = 0; i < np; i++) b.w[i] = wg[i];
  return FEM_OK;
}

int make_material(double E, double nu, double t, Material &m) {
  if (!(E > 0) || !(nu > -1.0 && nu < 0.5) || !(t > 0)) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "material E=%g nu=%g t=%g out of range", E, nu, t);
  }
  // Plane stress: mu = E / (2 (1 + nu)), lambda = 2 mu nu / (1 - nu)  (FEKernels.h:300-301)
  m.mu = 0.5 * E / (1.0 + nu);
  m.lambda = 2 * m.mu * nu / (1.0 - nu);
  m.t = t;
  return FEM_OK;
}

// Records events around `launch` when timing is requested.
template <typename Fn>
int timed(hipStream_t stream, double *elapsedSeconds, Fn launch) {
  hipEvent_t start = nullptr, stop = nullptr;
  if (elapsedSeconds) {
    FEM_HIP_CHECK(hipEventCreate(&start));
    FEM_HIP_CHECK(hipEventCreate(&stop));
    FEM_HIP_CHECK(hipEventRecord(start, stream));
  }
  int rc = launch();
  if (rc == FEM_OK) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = set_error(FEM_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  }
  if (elapsedSeconds) {
    if (rc == FEM_OK) {
      hipError_t e = hipEventRecord(stop, stream);
      if (e == hipSuccess) e = hipEventSynchronize(stop);
      float ms = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, start, stop);
      if (e != hipSuccess) rc = set_error(FEM_ERR_HIP, "timing events failed: %s", hipGetErrorString(e));
      *elapsedSeconds = double(ms) * 1e-3;
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
  }
  return rc;
}

constexpr int elem_waves(int order) { return order <= 2 ? 4 : 2; }

template <int P, typename T, bool RES, bool JAC, bool ATOMIC>
void launch_element(const ElemArgs &a, hipStream_t stream) {
  constexpr int W = elem_waves(P);
  constexpr int per = W * Ord<P>::EPW;
  const int grid = (a.count + per - 1) / per;
  if (grid > 0) element_kernel<P, T, W, RES, JAC, ATOMIC><<<grid, W * 64, 0, stream>>>(a);
}

template <int P, bool RES, bool JAC, bool ATOMIC>
void launch_element_prec(const ElemArgs &a, int precision, hipStream_t stream) {
  if (precision == FEM_PREC_MIXED) {
    launch_element<P, float, RES, JAC, ATOMIC>(a, stream);
  } else {
    launch_element<P, double, RES, JAC, ATOMIC>(a, stream);
  }
}

template <bool RES, bool JAC, bool ATOMIC>
void launch_element_order(int order, const ElemArgs &a, int precision, hipStream_t stream) {
  switch (order) {
    case 1: launch_element_prec<1, RES, JAC, ATOMIC>(a, precision, stream); break;
    case 2: launch_element_prec<2, RES, JAC, ATOMIC>(a, precision, stream); break;
    case 3: launch_element_prec<3, RES, JAC, ATOMIC>(a, precision, stream); break;
    default: launch_element_prec<4, RES, JAC, ATOMIC>(a, precision, stream); break;
  }
}

// Wavefronts per workgroup of the element-once kernels (order-4 annulus, FP64: 2 waves 3.13 ms, 4 waves
// 2.95 ms, 8 waves 3.05 ms; the workgroup's basis tables are formed once for more elements).
constexpr int kTiledWaves = 4;

template <int P, typename T, bool RES>
void launch_element_split(const SplitJob &job, const Basis1D &basis, int waves, hipStream_t stream) {
  auto go = [&](auto W) {
    constexpr int WAVES = decltype(W)::value;
    constexpr int per = WAVES * Ord<P>::EPW;
    const int grid = (job.count + per - 1) / per;
    if (grid > 0) element_tiled_kernel<P, T, WAVES, RES><<<grid, WAVES * 64, 0, stream>>>(job, basis);
  };
  if (job.accumulate) {
    // accumulate mode (a compile-time variant, so the default emission is unchanged), default shape
    constexpr int per = kTiledWaves * Ord<P>::EPW;
    const int grid = (job.count + per - 1) / per;
    if (grid > 0) element_tiled_kernel<P, T, kTiledWaves, RES, true><<<grid, kTiledWaves * 64, 0, stream>>>(job, basis);
    return;
  }
  // wavefronts per workgroup: the plan's (autotuned) choice
  if (waves == 1) go(std::integral_constant<int, 1>{});
  else if (waves == 2) go(std::integral_constant<int, 2>{});
  else go(std::integral_constant<int, kTiledWaves>{});
}

// Second pass of the element-once assembly: output dst[i] (C doubles from C * dst[i]) is the sum of the
// staged contributions in slots ptr[i] .. ptr[i+1]-1, in slot order; zero if there are none. Lane t covers
// component t % C of output t / C, so a wavefront reads consecutive slots.
template <int C>
__device__ __forceinline__ void gather_one(const int *dst, const int *ptr, const double *stage, double *out, int64_t i,
                                           int c, bool accumulate = false) {
  const int k0 = ptr[i], k1 = ptr[i + 1];
  double sum = 0;
  for (int k = k0; k < k1; k++) sum += __builtin_nontemporal_load(stage + int64_t(C) * k + c);
  double *o = out + int64_t(C) * dst[i] + c;
  if (accumulate) sum += *o;
  __builtin_nontemporal_store(sum, o);
}

template <int C>
__global__ __launch_bounds__(256) void gather_shared(const int *dst, const int *ptr, const double *stage,
                                                     double *out, int n, int accumulate) {
  const int64_t total = int64_t(n) * C, stride = int64_t(gridDim.x) * 256;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total; t += stride) {
    gather_one<C>(dst, ptr, stage, out, t / C, int(t % C), accumulate != 0);
  }
}

template <int C>
void launch_gather_shared(const int *dst, const int *ptr, const double *stage, double *out, int n, int accumulate,
                          hipStream_t stream) {
  const int64_t blocks = (int64_t(n) * C + 255) / 256;
  if (n > 0) {
    gather_shared<C><<<unsigned(std::min<int64_t>(blocks, 1 << 20)), 256, 0, stream>>>(dst, ptr, stage, out, n,
                                                                                        accumulate);
  }
}

// The gathers of a batch's element-once jobs: C = 4, the staged blocks (outStart[j]: job j's first output
// over all jobs, counts[2] in total), or C = 2, the staged residual entries (counts[3]; jobs without a
// residual have none).
template <int C>
__global__ __launch_bounds__(256) void gather_batch(const SplitJob *jobs, const int *outStart, const int *counts) {
  const int numJobs = counts[0];
  const int64_t total = int64_t(counts[C == 4 ? 2 : 3]) * C, stride = int64_t(gridDim.x) * 256;
  for (int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x; t < total; t += stride) {
    const int64_t i = t / C;
    const int j = find_job(outStart, numJobs, i);
    const SplitJob &job = jobs[j];
    if (C == 4) gather_one<4>(job.gdst, job.gptr, job.stage, job.values, i - outStart[j], int(t % C));
    else gather_one<2>(job.rdst, job.rptr, job.rstage, job.residual, i - outStart[j], int(t % C));
  }
}

// Element-once assembly of one plan: the element kernel stores exclusive outputs and stages the rest, then
// the gathers store the staged sums.
template <int P>
void launch_split(const SplitJob &job, const Basis1D &basis, int precision, int waves, hipStream_t s) {
  const bool res = job.residual != nullptr;
  if (precision == FEM_PREC_MIXED) {
    if (res) launch_element_split<P, float, true>(job, basis, waves, s);
    else launch_element_split<P, float, false>(job, basis, waves, s);
  } else {
    if (res) launch_element_split<P, double, true>(job, basis, waves, s);
    else launch_element_split<P, double, false>(job, basis, waves, s);
  }
  launch_gather_shared<4>(job.gdst, job.gptr, job.stage, job.values, job.numGather, job.accumulate, s);
  if (res) launch_gather_shared<2>(job.rdst, job.rptr, job.rstage, job.residual, job.numRgather, job.accumulate, s);
}

template <int P, bool RES, bool JAC>
void launch_element_nl(const ElemArgs &a, hipStream_t stream) {
  constexpr int W = elem_waves(P);
  constexpr int per = W * Ord<P>::EPW;
  const int grid = (a.count + per - 1) / per;
  if (grid > 0) element_kernel_nl<P, W, RES, JAC><<<grid, W * 64, 0, stream>>>(a);
}

template <bool RES, bool JAC>
void launch_element_nl_order(int order, const ElemArgs &a, hipStream_t stream) {
  switch (order) {
    case 1: launch_element_nl<1, RES, JAC>(a, stream); break;
    case 2: launch_element_nl<2, RES, JAC>(a, stream); break;
    case 3: launch_element_nl<3, RES, JAC>(a, stream); break;
    default: launch_element_nl<4, RES, JAC>(a, stream); break;
  }
}

// One launch group: jobs of one order with the same outputs.
struct OwnerLaunch {
  OwnerArgs args;
  int precision;
  bool det;         // deterministic (colour-phased) sums: per-part kernel
  bool res, jac;    // outputs
  bool rOnlyParts;  // parts sized for residual-only assembly
  bool stream;      // every part is one batch per wavefront: streaming kernel
  int streamGrid;   // persistent workgroups (2 per CU, multiple of 8)
};

template <int P, typename T, int MAXR, bool RES, bool JAC, bool DET>
void launch_owner(const OwnerLaunch &L, hipStream_t stream) {
  constexpr OwnerConfig cfg = owner_config(P);
  if constexpr (!DET && MAXR <= cfg.waves * 64) {
    if (L.stream) {
      // streamGrid is two workgroups per CU; 8-wave workgroups fill a CU's LDS alone.
      const int persistent = cfg.waves >= 8 ? std::max(8, L.streamGrid / 16 * 8) : L.streamGrid;
      const int grid = std::min(persistent, ((L.args.totalParts + 7) / 8) * 8);
      owner_stream_kernel<P, T, cfg.waves, cfg.maxBlocks, MAXR, RES, JAC><<<grid, cfg.waves * 64, 0, stream>>>(
          L.args);
      return;
    }
  }
  const int grid = ((L.args.totalParts + 7) / 8) * 8;
  owner_kernel<P, T, cfg.waves, cfg.maxBlocks, MAXR, RES, JAC, DET><<<grid, cfg.waves * 64, 0, stream>>>(L.args);
}

template <int P, typename T, bool DET>
void launch_owner_mode(const OwnerLaunch &L, hipStream_t stream) {
  constexpr OwnerConfig cfg = owner_config(P);
  if (L.jac && L.res) {
    launch_owner<P, T, cfg.maxRows, true, true, DET>(L, stream);
  } else if (L.jac) {
    launch_owner<P, T, cfg.maxRows, false, true, DET>(L, stream);
  } else if (L.rOnlyParts) {
    launch_owner<P, T, kResidualOnlyMaxRows, true, false, DET>(L, stream);
  } else {
    // Residual from a plan whose parts were sized for the Jacobian: rows fit the smaller bound.
    launch_owner<P, T, cfg.maxRows, true, false, DET>(L, stream);
  }
}

template <int P>
void launch_owner_order(const OwnerLaunch &L, hipStream_t stream) {
  if (L.precision == FEM_PREC_MIXED) {
    if (L.det) launch_owner_mode<P, float, true>(L, stream);
    else launch_owner_mode<P, float, false>(L, stream);
  } else {
    if (L.det) launch_owner_mode<P, double, true>(L, stream);
    else launch_owner_mode<P, double, false>(L, stream);
  }
}

void launch_owner_any(int order, const OwnerLaunch &L, hipStream_t stream) {
  switch (order) {
    case 1: launch_owner_order<1>(L, stream); break;
    case 2: launch_owner_order<2>(L, stream); break;
    case 3: launch_owner_order<3>(L, stream); break;
    default: launch_owner_order<4>(L, stream); break;
  }
}

int owner_waves(int order) {
  switch (order) {
    case 1: return owner_config(1).waves;
    case 2: return owner_config(2).waves;
    case 3: return owner_config(3).waves;
    default: return owner_config(4).waves;
  }
}

// Whether a plan's parts all fit one batch per wavefront of the owner workgroup.
bool plan_streams(const fem_plan *p) {
  const bool rOnly = (p->flags & FEM_PLAN_RESIDUAL_ONLY) != 0;
  const bool det = (p->flags & FEM_PLAN_DETERMINISTIC) != 0;
  const int epw = 64 / nodes_per_element(p->order);
  return !rOnly && !det && p->numCUs > 0 && p->maxPartVisits <= owner_waves(p->order) * epw;
}

int stream_grid(int numCUs) { return ((2 * numCUs + 7) / 8) * 8; }

SplitJob make_split_job(const fem_plan *p, const double *states, const double *coords, double *residual,
                        double *values, const Material &m) {
  SplitJob j{};
  j.conn = p->dConn;
  j.code = p->dCode;
  j.rcode = p->dRcode;
  j.gdst = p->dGdst;
  j.gptr = p->dGptr;
  j.rdst = p->dRdst;
  j.rptr = p->dRptr;
  j.states = states;
  j.coords = coords;
  j.residual = residual;
  j.values = values;
  j.stage = p->dStage;
  j.rstage = p->dRstage;
  j.mat = m;
  j.count = p->numElements;
  j.numBlocks = int(p->numBlocks);
  j.numNodes = p->numNodes;
  j.numGather = p->numGather;
  j.numRgather = residual ? p->numRgather : 0;
  return j;
}

MeshJob make_job(const fem_plan *p, const double *states, const double *coords, double *residual, double *values,
                 const Material &m) {
  MeshJob j{};
  j.visitConn = p->dVisitConn;
  j.visitEid = p->dVisitEid;
  j.nodeLrow = p->dNodeLrow;
  j.rows = p->dRows;
  j.rowLbase = p->dRowLbase;
  j.runs = reinterpret_cast<const int4 *>(p->dRuns);
  j.partDesc = reinterpret_cast<const int4 *>(p->dPartDesc);
  j.rank = reinterpret_cast<const uint32_t *>(p->dRank);
  j.colourOff = p->dColourOff;
  j.states = states;
  j.coords = coords;
  j.residual = residual;
  j.values = values;
  j.mu = m.mu;
  j.lambda = m.lambda;
  j.t = m.t;
  j.numParts = p->numParts;
  j.numColours = p->numColours;
  return j;
}

int check_common(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                 const double *nodeCoords, int precision) {
  if (order < 1 || order > 4) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  if (numElements < 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "numElements %d < 0", numElements);
  if (precision != FEM_PREC_FP64 && precision != FEM_PREC_MIXED) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown precision mode %d", precision);
  }
  if (numElements > 0 && (!connPtr || !conn || !states || !nodeCoords)) {
    return set_error(FEM_ERR_NULL_POINTER, "connPtr/conn/states/nodeCoords is NULL");
  }
  return FEM_OK;
}

template <typename TPtr>
int upload(const std::vector<TPtr> &h, TPtr **d, int64_t &bytes) {
  const size_t n = std::max<size_t>(h.size(), 1) * sizeof(TPtr);
  FEM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(d), n));
  if (!h.empty()) FEM_HIP_CHECK(hipMemcpy(*d, h.data(), h.size() * sizeof(TPtr), hipMemcpyHostToDevice));
  bytes += int64_t(n);
  return FEM_OK;
}

int alloc_doubles(int64_t n, double **d, int64_t &bytes) {
  const size_t sz = size_t(std::max<int64_t>(n, 1)) * sizeof(double);
  FEM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(d), sz));
  bytes += int64_t(sz);
  return FEM_OK;
}

void free_plan_arrays(fem_plan *p) {
  void *ptrs[] = {p->dVisitConn, p->dVisitEid, p->dNodeLrow, p->dRows,  p->dRowLbase, p->dRuns,  p->dColourOff,
                  p->dPartDesc,  p->dRank,     p->dJob,      p->dJobStart, p->dConn,   p->dCode,  p->dRcode,
                  p->dGdst,      p->dGptr,     p->dRdst,     p->dRptr,   p->dStage,    p->dRstage};
  for (void *q : ptrs) {
    if (q) (void)hipFree(q);
  }
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" {

int fem_residual(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                 const double *nodeCoords, double E, double nu, double t, double *residual, int precision,
                 void *stream, double *elapsedSeconds) {
  int rc = check_common(order, connPtr, conn, numElements, states, nodeCoords, precision);
  if (rc) return rc;
  if (numElements > 0 && !residual) return set_error(FEM_ERR_NULL_POINTER, "residual is NULL");
  ElemArgs a{};
  a.connPtr = connPtr;
  a.conn = conn;
  a.elemList = nullptr;
  a.first = 0;
  a.count = numElements;
  a.states = states;
  a.coords = nodeCoords;
  a.residual = residual;
  if ((rc = make_basis(order, a.basis)) || (rc = make_material(E, nu, t, a.mat))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return timed(s, elapsedSeconds, [&]() {
    launch_element_order<true, false, true>(order, a, precision, s);
    return FEM_OK;
  });
}

int fem_jacobian(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                 const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                 double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds) {
  int rc = check_common(order, connPtr, conn, numElements, states, nodeCoords, precision);
  if (rc) return rc;
  if (numElements > 0 && (!elementBCSRMap || !matEntries)) {
    return set_error(FEM_ERR_NULL_POINTER, "elementBCSRMap/matEntries is NULL");
  }
  ElemArgs a{};
  a.connPtr = connPtr;
  a.conn = conn;
  a.elemList = nullptr;
  a.first = 0;
  a.count = numElements;
  a.states = states;
  a.coords = nodeCoords;
  a.map = elementBCSRMap;
  a.residual = residual;
  a.values = matEntries;
  if ((rc = make_basis(order, a.basis)) || (rc = make_material(E, nu, t, a.mat))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return timed(s, elapsedSeconds, [&]() {
    if (residual) launch_element_order<true, true, true>(order, a, precision, s);
    else launch_element_order<false, true, true>(order, a, precision, s);
    return FEM_OK;
  });
}

int fem_residual_nonlinear(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                           const double *nodeCoords, double E, double nu, double t, double *residual, int precision,
                           void *stream, double *elapsedSeconds) {
  int rc = check_common(order, connPtr, conn, numElements, states, nodeCoords, precision);
  if (rc) return rc;
  if (precision != FEM_PREC_FP64) return set_error(FEM_ERR_UNSUPPORTED, "the NONLINEAR strain runs in FP64 only");
  if (numElements > 0 && !residual) return set_error(FEM_ERR_NULL_POINTER, "residual is NULL");
  ElemArgs a{};
  a.connPtr = connPtr;
  a.conn = conn;
  a.count = numElements;
  a.states = states;
  a.coords = nodeCoords;
  a.residual = residual;
  if ((rc = make_basis(order, a.basis)) || (rc = make_material(E, nu, t, a.mat))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return timed(s, elapsedSeconds, [&]() {
    launch_element_nl_order<true, false>(order, a, s);
    return FEM_OK;
  });
}

int fem_jacobian_nonlinear(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                           const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                           double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds) {
  int rc = check_common(order, connPtr, conn, numElements, states, nodeCoords, precision);
  if (rc) return rc;
  if (precision != FEM_PREC_FP64) return set_error(FEM_ERR_UNSUPPORTED, "the NONLINEAR strain runs in FP64 only");
  if (numElements > 0 && (!elementBCSRMap || !matEntries)) {
    return set_error(FEM_ERR_NULL_POINTER, "elementBCSRMap/matEntries is NULL");
  }
  ElemArgs a{};
  a.connPtr = connPtr;
  a.conn = conn;
  a.count = numElements;
  a.states = states;
  a.coords = nodeCoords;
  a.map = elementBCSRMap;
  a.residual = residual;
  a.values = matEntries;
  if ((rc = make_basis(order, a.basis)) || (rc = make_material(E, nu, t, a.mat))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return timed(s, elapsedSeconds, [&]() {
    if (residual) launch_element_nl_order<true, true>(order, a, s);
    else launch_element_nl_order<false, true>(order, a, s);
    return FEM_OK;
  });
}

int fem_jacobian_coloured(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                          const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                          const int *colourElements, const int *colourPtr, int numColours, double *residual,
                          double *matEntries, int precision, void *stream, double *elapsedSeconds) {
  int rc = check_common(order, connPtr, conn, numElements, states, nodeCoords, precision);
  if (rc) return rc;
  if (numElements > 0 && (!elementBCSRMap || !matEntries || !colourElements || !colourPtr)) {
    return set_error(FEM_ERR_NULL_POINTER, "elementBCSRMap/matEntries/colourElements/colourPtr is NULL");
  }
  if (numColours < 0 || (numElements > 0 && numColours == 0)) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "numColours %d invalid", numColours);
  }
  for (int c = 0; c < numColours; c++) {
    if (colourPtr[c + 1] < colourPtr[c]) return set_error(FEM_ERR_INVALID_ARGUMENT, "colourPtr not ascending");
  }
  if (numColours > 0 && (colourPtr[0] != 0 || colourPtr[numColours] != numElements)) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "colourPtr must span 0..numElements");
  }
  ElemArgs a{};
  a.connPtr = connPtr;
  a.conn = conn;
  a.states = states;
  a.coords = nodeCoords;
  a.map = elementBCSRMap;
  a.residual = residual;
  a.values = matEntries;
  if ((rc = make_basis(order, a.basis)) || (rc = make_material(E, nu, t, a.mat))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return timed(s, elapsedSeconds, [&]() {
    for (int c = 0; c < numColours; c++) {
      a.elemList = colourElements + colourPtr[c];
      a.first = 0;
      a.count = colourPtr[c + 1] - colourPtr[c];
      if (residual) launch_element_order<true, true, false>(order, a, precision, s);
      else launch_element_order<false, true, false>(order, a, precision, s);
    }
    return FEM_OK;
  });
}

}  // extern "C"

namespace {

int create_plan(int order, int numNodes, int numElements, const int *connPtr, const int *conn, const int *rowp,
                const int *cols, const double *nodeCoordsHost, int flags, double partScale, fem_plan **plan) {
  *plan = nullptr;
  PlanHost h;
  int rc = build_plan_host(order, numNodes, numElements, connPtr, conn, rowp, cols, nodeCoordsHost,
                           flags & ~FEM_PLAN_AUTOTUNE, h, partScale);
  if (rc) return rc;
  fem_plan *p = new fem_plan();
  p->order = order;
  p->numNodes = numNodes;
  p->numElements = numElements;
  p->numParts = h.numParts;
  p->numColours = h.numColours;
  p->flags = flags;
  p->numBlocks = h.numBlocks;
  p->elementVisits = int64_t(h.visitEid.size());
  p->maxPartVisits = h.maxPartVisits;
  (void)hipGetDevice(&p->device);
  if (hipDeviceGetAttribute(&p->numCUs, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess) p->numCUs = 0;
  const std::vector<int> jobStart = {0, h.numParts};
  int64_t bytes = 0;
  if ((rc = upload(h.visitConn, &p->dVisitConn, bytes)) || (rc = upload(h.visitEid, &p->dVisitEid, bytes)) ||
      (rc = upload(h.nodeLrow, &p->dNodeLrow, bytes)) || (rc = upload(h.rows, &p->dRows, bytes)) ||
      (rc = upload(h.rowLbase, &p->dRowLbase, bytes)) || (rc = upload(h.runs, &p->dRuns, bytes)) ||
      (rc = upload(h.colourOff, &p->dColourOff, bytes)) || (rc = upload(h.partDesc, &p->dPartDesc, bytes)) ||
      (rc = upload(h.rank, &p->dRank, bytes)) || (rc = upload(jobStart, &p->dJobStart, bytes))) {
    free_plan_arrays(p);
    delete p;
    return rc;
  }
  // Orders 3-4: Jacobian assemblies traverse the elements once (the owner-computes parts visit each element
  // 3-4 times there). Most blocks of an element's K_e are formed by that element alone and are stored
  // directly; the rest are staged and summed by a gather pass (SplitHost). Without room for int codes the
  // plan keeps the owner-computes path (as FEM_PLAN_OWNER_ONLY plans do).
  if (order >= 3 && !(flags & (FEM_PLAN_RESIDUAL_ONLY | FEM_PLAN_OWNER_ONLY))) {
    const int N = (order + 1) * (order + 1);
    std::vector<int> c(size_t(numElements) * N);
    for (int e = 0; e < numElements; e++) std::copy(conn + connPtr[e], conn + connPtr[e] + N, c.begin() + size_t(e) * N);
    SplitHost sp;
    if (build_split_host(order, numNodes, numElements, c.data(), rowp, h.rank, h.numBlocks, sp) == FEM_OK) {
      h = PlanHost();  // the owner tables are on the device; release their host copies first
      if ((rc = upload(c, &p->dConn, bytes)) || (rc = upload(sp.code, &p->dCode, bytes)) ||
          (rc = upload(sp.rcode, &p->dRcode, bytes)) || (rc = upload(sp.dst, &p->dGdst, bytes)) ||
          (rc = upload(sp.ptr, &p->dGptr, bytes)) || (rc = upload(sp.rdst, &p->dRdst, bytes)) ||
          (rc = upload(sp.rptr, &p->dRptr, bytes)) || (rc = alloc_doubles(4 * sp.slots, &p->dStage, bytes)) ||
          (rc = alloc_doubles(2 * sp.rslots, &p->dRstage, bytes))) {
        free_plan_arrays(p);
        delete p;
        return rc;
      }
      p->numGather = int(sp.dst.size());
      p->numRgather = int(sp.rdst.size());
      p->numSlots = sp.slots;
      p->numRslots = sp.rslots;
    } else {
      clear_error();
    }
  }
  if (hipMalloc(reinterpret_cast<void **>(&p->dJob), sizeof(MeshJob)) != hipSuccess) {
    free_plan_arrays(p);
    delete p;
    return set_error(FEM_ERR_OUT_OF_MEMORY, "hipMalloc of the plan's job record failed");
  }
  bytes += int64_t(sizeof(MeshJob));
  static_assert(sizeof(MeshJob) <= sizeof(p->hostJob), "job record size");
  p->deviceBytes = bytes;
  p->partScale = partScale;
  p->flags = flags & ~FEM_PLAN_AUTOTUNE;
  *plan = p;
  return FEM_OK;
}

// Device buffers of one timing run, freed on scope exit.
struct TuneBuffers {
  double *coords = nullptr, *states = nullptr, *res = nullptr, *vals = nullptr;
  hipStream_t stream = nullptr;
  ~TuneBuffers() {
    if (coords) (void)hipFree(coords);
    if (states) (void)hipFree(states);
    if (res) (void)hipFree(res);
    if (vals) (void)hipFree(vals);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace

extern "C" {

int fem_plan_create(int order, int numNodes, int numElements, const int *connPtr, const int *conn, const int *rowp,
                    const int *cols, const double *nodeCoordsHost, int flags, fem_plan **plan) {
  if (!plan) return set_error(FEM_ERR_NULL_POINTER, "plan is NULL");
  *plan = nullptr;
  if (!(flags & FEM_PLAN_AUTOTUNE)) {
    return create_plan(order, numNodes, numElements, connPtr, conn, rowp, cols, nodeCoordsHost, flags, 1.0, plan);
  }
  if (flags & (FEM_PLAN_DETERMINISTIC | FEM_PLAN_RESIDUAL_ONLY)) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "FEM_PLAN_AUTOTUNE takes no other plan flag");
  }
  if (!nodeCoordsHost) return set_error(FEM_ERR_NULL_POINTER, "FEM_PLAN_AUTOTUNE needs nodeCoordsHost");
  // Timing inputs: the mesh's coordinates, zero states, outputs of the mesh's size.
  TuneBuffers tb;
  const size_t nb = size_t(rowp[numNodes]);
  FEM_HIP_CHECK(hipStreamCreateWithFlags(&tb.stream, hipStreamNonBlocking));
  FEM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&tb.coords), size_t(numNodes) * 16));
  FEM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&tb.states), size_t(numNodes) * 16));
  FEM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&tb.res), size_t(numNodes) * 16));
  FEM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&tb.vals), nb * 32));
  FEM_HIP_CHECK(hipMemcpy(tb.coords, nodeCoordsHost, size_t(numNodes) * 16, hipMemcpyHostToDevice));
  FEM_HIP_CHECK(hipMemset(tb.states, 0, size_t(numNodes) * 16));
  fem_plan *best = nullptr;
  double bestSeconds = 0;
  // Orders 3-4 assemble their Jacobians element by element, independent of the part size; only full-size
  // parts are built there.
  const std::vector<double> scales = order >= 3 ? std::vector<double>{1.0} : std::vector<double>{1.0, 0.5, 0.25};
  for (const double scale : scales) {
    fem_plan *cand = nullptr;
    int rc = create_plan(order, numNodes, numElements, connPtr, conn, rowp, cols, nodeCoordsHost, flags, scale,
                         &cand);
    if (rc) {
      fem_plan_destroy(best);
      return rc;
    }
    double sec = 0, secMin = 1e30;
    for (int rep = 0; rep < 3 && !rc; rep++) {  // the first run warms up
      rc = fem_plan_assemble(cand, tb.states, tb.coords, 1.0, 0.3, 1.0, tb.res, tb.vals, FEM_PREC_MIXED, tb.stream,
                             &sec);
      if (rep > 0) secMin = std::min(secMin, sec);
    }
    if (rc) {
      fem_plan_destroy(cand);
      fem_plan_destroy(best);
      return rc;
    }
    if (order >= 3 && cand->dConn) {
      // orders 3-4: the element-once kernel's wavefronts per workgroup (lanes per element and elements per
      // wavefront are fixed by the order: N and 64 / N)
      int bestWaves = cand->tiledWaves;
      for (const int waves : {1, 2, kTiledWaves}) {
        cand->tiledWaves = waves;
        double s2 = 0, s2Min = 1e30;
        for (int rep = 0; rep < 3 && !rc; rep++) {
          rc = fem_plan_assemble(cand, tb.states, tb.coords, 1.0, 0.3, 1.0, tb.res, tb.vals, FEM_PREC_MIXED,
                                 tb.stream, &s2);
          if (rep > 0) s2Min = std::min(s2Min, s2);
        }
        if (rc) break;
        if (s2Min < secMin) {
          secMin = s2Min;
          bestWaves = waves;
        }
      }
      cand->tiledWaves = bestWaves;
      if (rc) {
        fem_plan_destroy(cand);
        fem_plan_destroy(best);
        return rc;
      }
    }
    cand->tunedSeconds = secMin;
    if (!best || secMin < bestSeconds) {
      fem_plan_destroy(best);
      best = cand;
      bestSeconds = secMin;
    } else {
      fem_plan_destroy(cand);
    }
  }
  *plan = best;
  return FEM_OK;
}

int fem_plan_tuning(const fem_plan *plan, double *partScale, double *seconds) {
  if (!plan) return set_error(FEM_ERR_NULL_POINTER, "plan is NULL");
  if (partScale) *partScale = plan->partScale;
  if (seconds) *seconds = plan->tunedSeconds;
  return FEM_OK;
}

int fem_plan_tuning_shape(const fem_plan *plan, int *wavesPerGroup, int *elementsPerWave, int *lanesPerElement) {
  if (!plan) return set_error(FEM_ERR_NULL_POINTER, "plan is NULL");
  const int N = nodes_per_element(plan->order);
  const bool tiled = plan->dConn != nullptr;  // orders 3-4 element-once Jacobians
  if (wavesPerGroup) *wavesPerGroup = tiled ? plan->tiledWaves : owner_waves(plan->order);
  if (elementsPerWave) *elementsPerWave = 64 / N;
  if (lanesPerElement) *lanesPerElement = N;
  return FEM_OK;
}

int fem_plan_info(const fem_plan *plan, int64_t *numParts, int64_t *numBlocks, int64_t *elementVisits,
                  int64_t *deviceBytes) {
  if (!plan) return set_error(FEM_ERR_NULL_POINTER, "plan is NULL");
  if (numParts) *numParts = plan->numParts;
  if (numBlocks) *numBlocks = plan->numBlocks;
  if (elementVisits) *elementVisits = plan->elementVisits;
  if (deviceBytes) *deviceBytes = plan->deviceBytes;
  return FEM_OK;
}

int fem_plan_assemble(fem_plan *plan, const double *states, const double *nodeCoords, double E, double nu, double t,
                      double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds) {
  return fem::plan_assemble(plan, states, nodeCoords, E, nu, t, residual, matEntries, precision, stream,
                            elapsedSeconds, false);
}

int fem_plan_assemble_accumulate(fem_plan *plan, const double *states, const double *nodeCoords, double E, double nu,
                                 double t, double *residual, double *matEntries, int precision, void *stream,
                                 double *elapsedSeconds) {
  return fem::plan_assemble(plan, states, nodeCoords, E, nu, t, residual, matEntries, precision, stream,
                            elapsedSeconds, true);
}

}  // extern "C"

int fem::plan_assemble(fem_plan *plan, const double *states, const double *nodeCoords, double E, double nu, double t,
                       double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds,
                       bool accumulate) {
  if (!plan) return set_error(FEM_ERR_NULL_POINTER, "plan is NULL");
  if (!states || !nodeCoords) return set_error(FEM_ERR_NULL_POINTER, "states/nodeCoords is NULL");
  if (!residual && !matEntries) return set_error(FEM_ERR_NULL_POINTER, "residual and matEntries are both NULL");
  if (precision != FEM_PREC_FP64 && precision != FEM_PREC_MIXED) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown precision mode %d", precision);
  }
  const bool rOnly = (plan->flags & FEM_PLAN_RESIDUAL_ONLY) != 0;
  if (rOnly && matEntries) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "plan was built residual-only; it cannot assemble the Jacobian");
  }
  Material m;
  int rc = make_material(E, nu, t, m);
  if (rc) return rc;
  OwnerLaunch L{};
  if ((rc = make_basis(plan->order, L.args.basis))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // The device job record is rewritten only when the request's pointers or material change.
  MeshJob job = make_job(plan, states, nodeCoords, residual, matEntries, m);
  job.accumulate = accumulate ? 1 : 0;
  if (!plan->jobUploaded || std::memcmp(&job, plan->hostJob, sizeof(MeshJob)) != 0) {
    FEM_HIP_CHECK(hipMemcpyAsync(plan->dJob, &job, sizeof(MeshJob), hipMemcpyHostToDevice, s));
    std::memcpy(plan->hostJob, &job, sizeof(MeshJob));
    plan->jobUploaded = true;
  }
  L.args.jobs = reinterpret_cast<const MeshJob *>(plan->dJob);
  L.args.jobStart = plan->dJobStart;
  L.args.numJobs = 1;
  L.args.totalParts = plan->numParts;
  L.precision = precision;
  L.det = (plan->flags & FEM_PLAN_DETERMINISTIC) != 0;
  L.res = residual != nullptr;
  L.jac = matEntries != nullptr;
  L.rOnlyParts = rOnly;
  L.stream = plan_streams(plan);
  L.streamGrid = stream_grid(plan->numCUs);
  if (plan->dConn && matEntries) {
    // Element-once assembly (orders 3-4): exclusive outputs stored by their element, the rest staged and
    // gathered. Deterministic, so it also serves FEM_PLAN_DETERMINISTIC plans.
    SplitJob job = make_split_job(plan, states, nodeCoords, residual, matEntries, m);
    job.accumulate = accumulate ? 1 : 0;
    return timed(s, elapsedSeconds, [&]() {
      if (plan->order == 3) launch_split<3>(job, L.args.basis, precision, plan->tiledWaves, s);
      else launch_split<4>(job, L.args.basis, precision, plan->tiledWaves, s);
      return FEM_OK;
    });
  }
  return timed(s, elapsedSeconds, [&]() {
    launch_owner_any(plan->order, L, s);
    return FEM_OK;
  });
}

extern "C" {

int fem_plan_destroy(fem_plan *plan) {
  if (!plan) return FEM_OK;
  free_plan_arrays(plan);
  delete plan;
  return FEM_OK;
}

// ---- batches of resident plans ------------------------------------------------------------------
}  // extern "C"

// 12 owner launch groups: (order - 1) * 3 + outputs, outputs 0: residual, 1: values, 2: both; and 4
// element-once groups (Jacobian jobs of orders 3-4 whose plans have element-once tables): (order - 3) * 2 +
// (1 if with residual).
struct fem_batch {
  struct Group {
    std::vector<MeshJob> jobs;
    std::vector<int> jobStart{0};
    bool stream = true;
  };
  struct SplitGroup {
    std::vector<SplitJob> jobs;
    std::vector<int> unitStart{0}, gStart{0}, rStart{0};  // prefix sums: units, staged blocks, staged residuals
    std::vector<int64_t> stageOff, rstageOff;             // each job's staging, in doubles into the arena
  };
  static constexpr int kGroups = 12;
  static constexpr int kSplitGroups = 4;
  Group groups[kGroups];
  SplitGroup split[kSplitGroups];
  // Staging of the element-once jobs: the batch's own (so that batches holding one plan can run
  // concurrently), grown at assembly to the jobs' total and kept.
  double *dArena = nullptr;
  int64_t arenaDoubles = 0, splitDoubles = 0;
  int maxJobs = 0;
  int flags = 0;
  int numCUs = 0;
  int numJobs = 0;
  // One device block at a fixed address (a captured graph refers to it), laid out as
  //   counts [kGroups + kSplitGroups][4] int (owner: numJobs, totalParts; element-once: numJobs, units,
  //     staged blocks, staged residuals)
  //   | jobStart [kGroups][maxJobs + 1] int | jobs [kGroups][maxJobs] MeshJob
  //   | split starts [kSplitGroups][3][maxJobs + 1] int (unitStart, gStart, rStart)
  //   | split jobs [kSplitGroups][maxJobs] SplitJob
  // and mirrored in pinned host memory, so each assembly stages its tables with one asynchronous copy.
  unsigned char *dTables = nullptr;
  unsigned char *hTables = nullptr;
  size_t jobStartOffset = 0, jobsOffset = 0, splitStartOffset = 0, splitJobsOffset = 0, tableBytes = 0;
  hipEvent_t staged = nullptr;  // recorded after the copy: the host mirror is rewritten only once it has run
  bool stagedPending = false;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // per precision
  // Jobs of deterministic and residual-only plans (batches without graphs only): each runs as its plan's own
  // assembly on the batch's stream, after the groups.
  struct Solo {
    fem_plan *plan;
    const double *states, *coords;
    double E, nu, t;
    double *residual, *values;
  };
  std::vector<Solo> solo;
  bool assembled = false;                          // fem_batch_assemble has staged job tables
  hipStream_t captureStream = nullptr;             // graphs are recorded here and replayed on any stream

  int *dCounts() const { return reinterpret_cast<int *>(dTables); }
  int *dJobStart() const { return reinterpret_cast<int *>(dTables + jobStartOffset); }
  MeshJob *dJobs() const { return reinterpret_cast<MeshJob *>(dTables + jobsOffset); }
  int *hCounts() const { return reinterpret_cast<int *>(hTables); }
  int *hJobStart() const { return reinterpret_cast<int *>(hTables + jobStartOffset); }
  MeshJob *hJobs() const { return reinterpret_cast<MeshJob *>(hTables + jobsOffset); }
  // element-once group g: start array k (0 units, 1 staged blocks, 2 staged residuals) and jobs
  size_t splitStartAt(int g, int k) const { return splitStartOffset + (size_t(g) * 3 + k) * (maxJobs + 1) * sizeof(int); }
  size_t splitJobsAt(int g) const { return splitJobsOffset + size_t(g) * maxJobs * sizeof(SplitJob); }
};

namespace {

void free_batch(fem_batch *b) {
  for (auto &g : b->graph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  if (b->staged) {
    if (b->stagedPending) (void)hipEventSynchronize(b->staged);
    (void)hipEventDestroy(b->staged);
  }
  if (b->captureStream) (void)hipStreamDestroy(b->captureStream);
  if (b->dArena) (void)hipFree(b->dArena);
  if (b->dTables) (void)hipFree(b->dTables);
  if (b->hTables) (void)hipHostFree(b->hTables);
}

OwnerLaunch batch_group_launch(const fem_batch *b, int gi, int precision, bool fromDevice) {
  const fem_batch::Group &g = b->groups[gi];
  OwnerLaunch L{};
  make_basis(gi / 3 + 1, L.args.basis);
  L.args.jobs = b->dJobs() + size_t(gi) * b->maxJobs;
  L.args.jobStart = b->dJobStart() + size_t(gi) * (b->maxJobs + 1);
  L.args.counts = fromDevice ? b->dCounts() + 4 * gi : nullptr;
  L.args.numJobs = int(g.jobs.size());
  L.args.totalParts = g.jobStart.back();
  L.precision = precision;
  L.det = false;
  L.res = gi % 3 != 1;
  L.jac = gi % 3 != 0;
  L.rOnlyParts = false;
  L.stream = fromDevice || (g.stream && b->numCUs > 0);
  L.streamGrid = stream_grid(b->numCUs);
  return L;
}

// Launches of element-once group g (order 3 + g / 2; with residual if g is odd): the tiled kernel over the
// group's units, then the gathers of its staged blocks and residual entries. The kernels read the group's
// counts from device memory; `fixedGrid` sizes the grids for any batch (graph recording), otherwise from the
// host's counts.
template <int P, typename T, bool RES>
void launch_split_group_t(const fem_batch *b, int g, bool fixedGrid, hipStream_t s) {
  const fem_batch::SplitGroup &sg = b->split[g];
  const int *counts = b->dCounts() + 4 * (fem_batch::kGroups + g);
  const SplitJob *jobs = reinterpret_cast<const SplitJob *>(b->dTables + b->splitJobsAt(g));
  const int *unitStart = reinterpret_cast<const int *>(b->dTables + b->splitStartAt(g, 0));
  const int *gStart = reinterpret_cast<const int *>(b->dTables + b->splitStartAt(g, 1));
  const int *rStart = reinterpret_cast<const int *>(b->dTables + b->splitStartAt(g, 2));
  const int cap = std::max(8 * b->numCUs, 64);
  auto grid = [&](int64_t work) { return unsigned(fixedGrid ? cap : std::max<int64_t>(1, std::min<int64_t>(work, cap))); };
  Basis1D basis;
  make_basis(P, basis);
  element_tiled_batch<P, T, kTiledWaves, RES><<<grid(sg.unitStart.back()), kTiledWaves * 64, 0, s>>>(jobs, unitStart,
                                                                                                   counts, basis);
  gather_batch<4><<<grid((int64_t(sg.gStart.back()) * 4 + 255) / 256), 256, 0, s>>>(jobs, gStart, counts);
  if (RES) gather_batch<2><<<grid((int64_t(sg.rStart.back()) * 2 + 255) / 256), 256, 0, s>>>(jobs, rStart, counts);
}

void launch_split_group(const fem_batch *b, int g, int precision, bool fixedGrid, hipStream_t s) {
  const bool mixed = precision == FEM_PREC_MIXED;
  switch (g) {
    case 0: mixed ? launch_split_group_t<3, float, false>(b, g, fixedGrid, s) : launch_split_group_t<3, double, false>(b, g, fixedGrid, s); break;
    case 1: mixed ? launch_split_group_t<3, float, true>(b, g, fixedGrid, s) : launch_split_group_t<3, double, true>(b, g, fixedGrid, s); break;
    case 2: mixed ? launch_split_group_t<4, float, false>(b, g, fixedGrid, s) : launch_split_group_t<4, double, false>(b, g, fixedGrid, s); break;
    default: mixed ? launch_split_group_t<4, float, true>(b, g, fixedGrid, s) : launch_split_group_t<4, double, true>(b, g, fixedGrid, s); break;
  }
}

// Graph of the 12 group launches, each reading its job count from device memory (empty groups exit).
// Recorded on the batch's own stream (capture is not allowed on the null stream); it is replayed on the
// caller's stream.
int capture_batch_graph(fem_batch *b, int precision) {
  if (!b->captureStream) FEM_HIP_CHECK(hipStreamCreateWithFlags(&b->captureStream, hipStreamNonBlocking));
  hipStream_t s = b->captureStream;
  hipGraph_t graph = nullptr;
  FEM_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int gi = 0; gi < fem_batch::kGroups; gi++) {
    OwnerLaunch L = batch_group_launch(b, gi, precision, true);
    L.args.totalParts = L.streamGrid;  // host-side grid bound only; the kernel reads the device count
    launch_owner_any(gi / 3 + 1, L, s);
  }
  for (int g = 0; g < fem_batch::kSplitGroups; g++) launch_split_group(b, g, precision, true, s);
  FEM_HIP_CHECK(hipStreamEndCapture(s, &graph));
  hipError_t e = hipGraphInstantiate(&b->graph[precision], graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return set_error(FEM_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  return FEM_OK;
}

}  // namespace

int fem::batch_prepare_graph(fem_batch *b, int precision) {
  if (!b) return set_error(FEM_ERR_NULL_POINTER, "batch is NULL");
  if (precision != FEM_PREC_FP64 && precision != FEM_PREC_MIXED) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown precision mode %d", precision);
  }
  if (!(b->flags & FEM_BATCH_GRAPH) || b->graph[precision]) return FEM_OK;
  int rc = capture_batch_graph(b, precision);
  if (rc) return rc;
  if (!b->assembled) {
    // One replay while every job count in the tables is still zero (all kernels exit at once): the runtime's
    // first-launch allocations (scratch of kernels that spill, graph kernel arguments) happen now, not later.
    FEM_HIP_CHECK(hipGraphLaunch(b->graph[precision], b->captureStream));
    FEM_HIP_CHECK(hipStreamSynchronize(b->captureStream));
  }
  return FEM_OK;
}

extern "C" {

int fem_batch_create(int maxJobs, int flags, fem_batch **batch) {
  if (!batch) return set_error(FEM_ERR_NULL_POINTER, "batch is NULL");
  if (maxJobs < 1) return set_error(FEM_ERR_INVALID_ARGUMENT, "maxJobs %d < 1", maxJobs);
  fem_batch *b = new fem_batch();
  b->maxJobs = maxJobs;
  b->flags = flags;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&b->numCUs, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) b->numCUs = 0;
  const size_t G = fem_batch::kGroups, SG = fem_batch::kSplitGroups;
  b->jobStartOffset = 4 * (G + SG) * sizeof(int);
  b->jobsOffset = (b->jobStartOffset + G * size_t(maxJobs + 1) * sizeof(int) + 255) / 256 * 256;
  b->splitStartOffset = b->jobsOffset + G * size_t(maxJobs) * sizeof(MeshJob);
  b->splitJobsOffset = (b->splitStartOffset + SG * 3 * size_t(maxJobs + 1) * sizeof(int) + 255) / 256 * 256;
  b->tableBytes = b->splitJobsOffset + SG * size_t(maxJobs) * sizeof(SplitJob);
  if (hipMalloc(reinterpret_cast<void **>(&b->dTables), b->tableBytes) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&b->hTables), b->tableBytes, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess) {
    free_batch(b);
    delete b;
    return set_error(FEM_ERR_OUT_OF_MEMORY, "allocating batch job tables (%zu bytes) failed", b->tableBytes);
  }
  std::memset(b->hTables, 0, b->tableBytes);
  if (hipMemcpy(b->dTables, b->hTables, b->tableBytes, hipMemcpyHostToDevice) != hipSuccess) {
    free_batch(b);
    delete b;
    return set_error(FEM_ERR_HIP, "initialising batch job tables failed");
  }
  *batch = b;
  return FEM_OK;
}

int fem_batch_clear(fem_batch *batch) {
  if (!batch) return set_error(FEM_ERR_NULL_POINTER, "batch is NULL");
  for (auto &g : batch->groups) {
    g.jobs.clear();
    g.jobStart.assign(1, 0);
    g.stream = true;
  }
  for (auto &g : batch->split) {
    g.jobs.clear();
    g.unitStart.assign(1, 0);
    g.gStart.assign(1, 0);
    g.rStart.assign(1, 0);
    g.stageOff.clear();
    g.rstageOff.clear();
  }
  batch->solo.clear();
  batch->splitDoubles = 0;
  batch->numJobs = 0;
  return FEM_OK;
}

int fem_batch_add(fem_batch *batch, fem_plan *plan, const double *states, const double *nodeCoords, double E,
                  double nu, double t, double *residual, double *matEntries) {
  return fem::batch_add_staged(batch, plan, states, nodeCoords, E, nu, t, residual, matEntries, nullptr);
}

}  // extern "C"

namespace fem {

int64_t batch_stage_doubles(const fem_plan *plan, bool withResidual) {
  if (!plan || !plan->dConn) return 0;
  return 4 * plan->numSlots + (withResidual ? 2 * plan->numRslots : 0);
}

bool plan_streams_public(const fem_plan *plan) { return plan && plan_streams(plan); }

int batch_add_staged(fem_batch *batch, fem_plan *plan, const double *states, const double *nodeCoords, double E,
                     double nu, double t, double *residual, double *matEntries, double *stage) {
  if (!batch || !plan) return set_error(FEM_ERR_NULL_POINTER, "batch/plan is NULL");
  if (!states || !nodeCoords) return set_error(FEM_ERR_NULL_POINTER, "states/nodeCoords is NULL");
  if (!residual && !matEntries) return set_error(FEM_ERR_NULL_POINTER, "residual and matEntries are both NULL");
  Material m;
  int rc = make_material(E, nu, t, m);
  if (rc) return rc;
  if ((plan->flags & FEM_PLAN_DETERMINISTIC) || (plan->flags & FEM_PLAN_RESIDUAL_ONLY)) {
    if (batch->flags & FEM_BATCH_GRAPH) {
      return set_error(FEM_ERR_UNSUPPORTED, "graph batches take plans built with default flags");
    }
    if ((plan->flags & FEM_PLAN_RESIDUAL_ONLY) && matEntries) {
      return set_error(FEM_ERR_INVALID_ARGUMENT, "plan was built residual-only; it cannot assemble the Jacobian");
    }
    if (int(batch->solo.size()) >= batch->maxJobs) {
      return set_error(FEM_ERR_INVALID_ARGUMENT, "batch group is full (%d jobs)", batch->maxJobs);
    }
    batch->solo.push_back(fem_batch::Solo{plan, states, nodeCoords, E, nu, t, residual, matEntries});
    batch->numJobs++;
    return FEM_OK;
  }
  // Jacobian jobs of plans with element-once tables take the element-once groups; their staging is placed
  // in the batch's arena at assembly.
  if (plan->dConn && matEntries) {
    fem_batch::SplitGroup &g = batch->split[(plan->order - 3) * 2 + (residual ? 1 : 0)];
    if (int(g.jobs.size()) >= batch->maxJobs) {
      return set_error(FEM_ERR_INVALID_ARGUMENT, "batch group is full (%d jobs)", batch->maxJobs);
    }
    const SplitJob job = make_split_job(plan, states, nodeCoords, residual, matEntries, m);
    const int per = kTiledWaves * (plan->order == 3 ? Ord<3>::EPW : Ord<4>::EPW);
    const int64_t units = g.unitStart.back() + (int64_t(plan->numElements) + per - 1) / per;
    const int64_t staged = int64_t(g.gStart.back()) + job.numGather;
    const int64_t rstaged = int64_t(g.rStart.back()) + job.numRgather;
    if (units > INT32_MAX || staged > INT32_MAX || rstaged > INT32_MAX) {
      return set_error(FEM_ERR_UNSUPPORTED, "batch group exceeds int work counts");
    }
    g.jobs.push_back(job);
    g.unitStart.push_back(int(units));
    g.gStart.push_back(int(staged));
    g.rStart.push_back(int(rstaged));
    if (stage) {
      // caller-owned staging: fixed now, never in the batch's arena
      g.jobs.back().stage = stage;
      g.jobs.back().rstage = stage + 4 * plan->numSlots;
      g.stageOff.push_back(-1);
      g.rstageOff.push_back(-1);
    } else {
      g.stageOff.push_back(batch->splitDoubles);
      batch->splitDoubles += 4 * plan->numSlots;
      g.rstageOff.push_back(batch->splitDoubles);
      if (residual) batch->splitDoubles += 2 * plan->numRslots;
    }
    batch->numJobs++;
    return FEM_OK;
  }
  if ((batch->flags & FEM_BATCH_GRAPH) && !plan_streams(plan)) {
    return set_error(FEM_ERR_UNSUPPORTED, "graph batches need plans whose parts are one batch per wavefront");
  }
  const int outputs = (residual && matEntries) ? 2 : (matEntries ? 1 : 0);
  fem_batch::Group &g = batch->groups[(plan->order - 1) * 3 + outputs];
  if (int(g.jobs.size()) >= batch->maxJobs) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "batch group is full (%d jobs)", batch->maxJobs);
  }
  g.jobs.push_back(make_job(plan, states, nodeCoords, residual, matEntries, m));
  g.jobStart.push_back(g.jobStart.back() + plan->numParts);
  g.stream = g.stream && plan_streams(plan);
  batch->numJobs++;
  return FEM_OK;
}

}  // namespace fem

extern "C" {

int fem_batch_assemble(fem_batch *batch, int precision, void *stream, double *elapsedSeconds) {
  if (!batch) return set_error(FEM_ERR_NULL_POINTER, "batch is NULL");
  if (precision != FEM_PREC_FP64 && precision != FEM_PREC_MIXED) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown precision mode %d", precision);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool useGraph = (batch->flags & FEM_BATCH_GRAPH) != 0;
  // The previous call's copy must have run before the pinned mirror is rewritten.
  if (batch->stagedPending) {
    FEM_HIP_CHECK(hipEventSynchronize(batch->staged));
    batch->stagedPending = false;
  }
  // The element-once jobs' staging arena: grown (after the device has drained, so no earlier assembly
  // still uses the old one) and never shrunk.
  if (batch->splitDoubles > batch->arenaDoubles) {
    // at least doubled, so that a stream of growing batches synchronises the device only a few times
    const int64_t want = std::max(batch->splitDoubles, 2 * batch->arenaDoubles);
    if (batch->dArena) {
      FEM_HIP_CHECK(hipDeviceSynchronize());
      FEM_HIP_CHECK(hipFree(batch->dArena));
    }
    batch->dArena = nullptr;
    batch->arenaDoubles = 0;
    for (const int64_t n : {want, batch->splitDoubles}) {
      if (hipMalloc(reinterpret_cast<void **>(&batch->dArena), size_t(n) * sizeof(double)) == hipSuccess) {
        batch->arenaDoubles = n;
        break;
      }
      (void)hipGetLastError();
      batch->dArena = nullptr;
    }
    if (!batch->dArena) {
      return set_error(FEM_ERR_OUT_OF_MEMORY, "allocating %lld doubles of batch staging failed",
                       (long long)batch->splitDoubles);
    }
  }
  for (auto &g : batch->split) {
    for (size_t j = 0; j < g.jobs.size(); j++) {
      if (g.stageOff[j] < 0) continue;  // caller-owned staging
      g.jobs[j].stage = batch->dArena + g.stageOff[j];
      g.jobs[j].rstage = batch->dArena + g.rstageOff[j];
    }
  }
  // Counts of all groups and job tables of non-empty groups, staged in the mirror and copied up to the
  // end of the last non-empty group's jobs.
  size_t copyEnd = batch->jobsOffset;
  for (int gi = 0; gi < fem_batch::kGroups; gi++) {
    fem_batch::Group &g = batch->groups[gi];
    batch->hCounts()[4 * gi] = int(g.jobs.size());
    batch->hCounts()[4 * gi + 1] = g.jobStart.back();
    if (g.jobs.empty()) continue;
    std::memcpy(batch->hJobs() + size_t(gi) * batch->maxJobs, g.jobs.data(), g.jobs.size() * sizeof(MeshJob));
    std::memcpy(batch->hJobStart() + size_t(gi) * (batch->maxJobs + 1), g.jobStart.data(),
                g.jobStart.size() * sizeof(int));
    copyEnd = batch->jobsOffset + (size_t(gi) * batch->maxJobs + g.jobs.size()) * sizeof(MeshJob);
  }
  for (int g = 0; g < fem_batch::kSplitGroups; g++) {
    const fem_batch::SplitGroup &sg = batch->split[g];
    int *cnt = batch->hCounts() + 4 * (fem_batch::kGroups + g);
    cnt[0] = int(sg.jobs.size());
    cnt[1] = sg.unitStart.back();
    cnt[2] = sg.gStart.back();
    cnt[3] = sg.rStart.back();
    if (sg.jobs.empty()) continue;
    const std::vector<int> *starts[3] = {&sg.unitStart, &sg.gStart, &sg.rStart};
    for (int k = 0; k < 3; k++) {
      std::memcpy(batch->hTables + batch->splitStartAt(g, k), starts[k]->data(), starts[k]->size() * sizeof(int));
    }
    std::memcpy(batch->hTables + batch->splitJobsAt(g), sg.jobs.data(), sg.jobs.size() * sizeof(SplitJob));
    copyEnd = batch->splitJobsAt(g) + sg.jobs.size() * sizeof(SplitJob);
  }
  FEM_HIP_CHECK(hipMemcpyAsync(batch->dTables, batch->hTables, copyEnd, hipMemcpyHostToDevice, s));
  FEM_HIP_CHECK(hipEventRecord(batch->staged, s));
  batch->stagedPending = true;
  batch->assembled = true;
  if (useGraph && !batch->graph[precision]) {
    int rc = capture_batch_graph(batch, precision);
    if (rc) return rc;
  }
  return timed(s, elapsedSeconds, [&]() -> int {
    if (useGraph) {
      hipError_t e = hipGraphLaunch(batch->graph[precision], s);
      if (e != hipSuccess) return set_error(FEM_ERR_HIP, "hipGraphLaunch failed: %s", hipGetErrorString(e));
      return FEM_OK;
    }
    for (int gi = 0; gi < fem_batch::kGroups; gi++) {
      fem_batch::Group &g = batch->groups[gi];
      if (g.jobs.empty()) continue;
      const OwnerLaunch L = batch_group_launch(batch, gi, precision, false);
      if (L.stream) {
        launch_owner_any(gi / 3 + 1, L, s);
      } else {
        // Parts too large for streaming: one per-part launch per job.
        for (size_t j = 0; j < g.jobs.size(); j++) {
          OwnerLaunch Lj = L;
          Lj.args.jobs = L.args.jobs + j;
          Lj.args.numJobs = 1;
          Lj.args.totalParts = g.jobs[j].numParts;
          launch_owner_any(gi / 3 + 1, Lj, s);
        }
      }
    }
    for (int g = 0; g < fem_batch::kSplitGroups; g++) {
      if (!batch->split[g].jobs.empty()) launch_split_group(batch, g, precision, false, s);
    }
    for (const fem_batch::Solo &j : batch->solo) {
      const int rc = fem::plan_assemble(j.plan, j.states, j.coords, j.E, j.nu, j.t, j.residual, j.values, precision, s,
                                        nullptr, false);
      if (rc) return rc;
    }
    return FEM_OK;
  });
}

int fem_batch_destroy(fem_batch *batch) {
  if (!batch) return FEM_OK;
  free_batch(batch);
  delete batch;
  return FEM_OK;
}

int fem_bsr_spmv(int numRows, const int *rowp, const int *cols, const double *values, const double *x, double *y,
                 void *stream) {
  if (numRows < 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "numRows < 0");
  if (numRows == 0) return FEM_OK;
  if (!rowp || !cols || !values || !x || !y) return set_error(FEM_ERR_NULL_POINTER, "NULL pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  bsr_spmv_kernel<<<(numRows + 3) / 4, 256, 0, s>>>(numRows, rowp, cols, values, x, y);
  FEM_HIP_CHECK(hipGetLastError());
  return FEM_OK;
}

}  // extern "C"
