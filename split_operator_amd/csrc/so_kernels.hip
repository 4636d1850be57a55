// so_kernels.hip — gfx950 kernels of the split-operator propagation step.
//
// One kernel family, `lane_pass`, executes a compiled pass (see so_device.hpp):
//   * a workgroup owns a tile of `tl` lanes along the pass axis; lanes are read
//     with the element order that keeps global accesses coalesced (lane index
//     fastest for strided lanes, element index fastest for contiguous lanes);
//   * element-wise stages (diagonal exp(-iHdt) tables, border masks, deferred
//     renormalisation factors, FP64 norm accumulation) are applied in registers
//     while the tile is loaded and stored;
//   * lane transforms run in LDS: Stockham autosort FFT with radix 8/4/2/3
//     passes and table twiddles (rounded from f64), or the lane matrix product
//     on MFMA (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64);
//   * norms are reduced per workgroup, written as partials, and the last
//     workgroup to finish sums them in a fixed order (deterministic) and runs
//     the pass's scalar program (LossChecker, LeakControl bookkeeping).
#include "so_device.hpp"

#include <math.h>

using namespace sod;

namespace {

template <typename T>
struct CT;
template <>
struct CT<float> {
  using T2 = float2;
};
template <>
struct CT<double> {
  using T2 = double2;
};

template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) {
  T2 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) {
  T2 r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
// multiply by exp(-i pi/2) = -i (forward) or +i (inverse)
template <bool INV, typename T2>
__device__ __forceinline__ T2 rot90(T2 z) {
  T2 r;
  if (INV) {
    r.x = -z.y;
    r.y = z.x;
  } else {
    r.x = z.y;
    r.y = -z.x;
  }
  return r;
}

// ---------------------------------------------------------------- small DFTs
template <bool INV, typename T2>
__device__ __forceinline__ void dft2(T2* v) {
  T2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <bool INV, typename T2>
__device__ __forceinline__ void dft4(T2* v) {
  T2 t0 = cadd(v[0], v[2]);
  T2 t1 = csub(v[0], v[2]);
  T2 t2 = cadd(v[1], v[3]);
  T2 t3 = rot90<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <bool INV, typename T, typename T2>
__device__ __forceinline__ void dft8(T2* v) {
  const T c = (T)0.70710678118654752440084436210485;
  T2 e[4] = {v[0], v[2], v[4], v[6]};
  T2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(e);
  dft4<INV>(o);
  T2 o1, o3;
  if (INV) {  // w8 = exp(+i pi/4)
    o1.x = c * (o[1].x - o[1].y);
    o1.y = c * (o[1].x + o[1].y);
    o3.x = -c * (o[3].x + o[3].y);
    o3.y = c * (o[3].x - o[3].y);
  } else {  // w8 = exp(-i pi/4)
    o1.x = c * (o[1].x + o[1].y);
    o1.y = c * (o[1].y - o[1].x);
    o3.x = c * (o[3].y - o[3].x);
    o3.y = -c * (o[3].x + o[3].y);
  }
  T2 o2 = rot90<INV>(o[2]);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <bool INV, typename T, typename T2>
__device__ __forceinline__ void dft3(T2* v) {
  const T s = (T)0.86602540378443864676372317075294;  // sin(2 pi / 3)
  T2 t = cadd(v[1], v[2]);
  T2 u = csub(v[1], v[2]);
  T2 m;
  m.x = v[0].x - (T)0.5 * t.x;
  m.y = v[0].y - (T)0.5 * t.y;
  T2 su;
  su.x = s * u.x;
  su.y = s * u.y;
  T2 r = rot90<INV>(su);
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int R, bool INV, typename T, typename T2>
__device__ __forceinline__ void dft(T2* v) {
  if constexpr (R == 2) dft2<INV>(v);
  if constexpr (R == 3) dft3<INV, T>(v);
  if constexpr (R == 4) dft4<INV>(v);
  if constexpr (R == 8) dft8<INV, T>(v);
}

// One Stockham radix-R pass over `tl` lanes of length n held in LDS (row pitch
// ld). Butterfly j of a lane reads x[j + r n/R], twiddles by W_{Ls R}^{r k}
// (k = j mod Ls), and writes y[(j/Ls) Ls R + k + r Ls]. All reads land in
// registers before the barrier, so the pass is done in place.
template <int R, bool INV, typename T, typename T2>
__device__ void radix_pass(T2* lds, int tl, int n, int ld, int Ls, const T2* __restrict__ tw, T scale,
                           bool last, int tid) {
  constexpr int MAXB = FFT_EPT / R;
  const int nb = n / R;
  const int total = tl * nb;
  const int twstep = n / (Ls * R);
  T2 v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = tid + q * NTHREADS;
    if (b < total) {
      const int l = b / nb;
      const int j = b - l * nb;
      const int k = j % Ls;
      const T2* src = lds + l * ld;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[j + r * nb];
      if (Ls > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          T2 w = tw[r * k * twstep];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft<R, INV, T>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = tid + q * NTHREADS;
    if (b < total) {
      const int l = b / nb;
      const int j = b - l * nb;
      const int k = j % Ls;
      T2* dst = lds + l * ld + (j / Ls) * Ls * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        T2 z = v[q][r];
        if (last) {
          z.x *= scale;
          z.y *= scale;
        }
        dst[r * Ls] = z;
      }
    }
  }
  __syncthreads();
}

template <bool INV, typename T, typename T2>
__device__ void fft_lds_dir(T2* lds, int tl, int n, int ld, const XformDesc& x, int tid) {
  const T2* tw = (const T2*)x.tw;
  const T scale = (T)x.scale;
  int Ls = 1;
  const int nrad = x.nrad;
  for (int p = 0; p < nrad; ++p) {
    const int R = x.rad[p];
    const bool last = (p == nrad - 1);
    switch (R) {
      case 8: radix_pass<8, INV, T>(lds, tl, n, ld, Ls, tw, scale, last, tid); break;
      case 4: radix_pass<4, INV, T>(lds, tl, n, ld, Ls, tw, scale, last, tid); break;
      case 3: radix_pass<3, INV, T>(lds, tl, n, ld, Ls, tw, scale, last, tid); break;
      default: radix_pass<2, INV, T>(lds, tl, n, ld, Ls, tw, scale, last, tid); break;
    }
    Ls *= R;
  }
}

template <typename T, typename T2>
__device__ void fft_lds(T2* lds, int tl, int n, int ld, const XformDesc& x, int tid) {
  if (x.inverse)
    fft_lds_dir<true, T>(lds, tl, n, ld, x, tid);
  else
    fft_lds_dir<false, T>(lds, tl, n, ld, x, tid);
}

// ---------------------------------------------------------------- MFMA lane matrix product
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Output rows held by lane group lk = L/16 in accumulator j: 4 lk + j for the
// f32 instruction, lk + 4 j for the f64 one (measured on gfx950 with exact
// small-integer operands).
template <>
struct Mfma<float> {
  using acc_t = v4f;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int j) { return 4 * lk + j; }
};
template <>
struct Mfma<double> {
  using acc_t = v4d;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int j) { return lk + 4 * j; }
};

// out-lanes = M . in-lanes for `tl` lanes in LDS. Real matrix: GEMM rows are
// (lane, re|im) pairs, K = n, columns = n. Complex matrix: rows are lanes seen
// as 2n reals, K = 2n, columns = 2n, B holds the real 2n x 2n form of M^T.
// 16x16x4 MFMA layout: lane L supplies A[L%16][L/16] and B[L/16][L%16] and
// holds D[Mfma<T>::row(L/16, j)][L%16], j = 0..3.
template <typename T, typename T2>
__device__ void mat_lds(const T2* in, T2* out, int tl, int ld, const XformDesc& x, int tid) {
  using M = Mfma<T>;
  const T* A = (const T*)in;
  T* Cm = (T*)out;
  const T* __restrict__ B = (const T*)x.B;
  const int rows = x.cplx ? tl : 2 * tl;
  const int rtiles = (rows + 15) / 16;
  const int ctiles = x.npad / 16;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int lr = lane & 15;
  const int lk = lane >> 4;
  const int kdim = x.kdim, kpad = x.kpad, npad = x.npad, ncols = x.ncols;
  for (int t = wave; t < rtiles * ctiles; t += NTHREADS / 64) {
    const int rt = t / ctiles;
    const int ct = t - rt * ctiles;
    const int row = rt * 16 + lr;
    const bool rv = row < rows;
    int abase, astride;
    if (x.cplx) {
      abase = row * ld * 2;
      astride = 1;
    } else {
      abase = (row >> 1) * ld * 2 + (row & 1);
      astride = 2;
    }
    typename M::acc_t acc = {0, 0, 0, 0};
    const T* bcol = B + ct * 16 + lr;
    for (int kk = 0; kk < kpad; kk += 4) {
      const int k = kk + lk;
      const T a = (rv && k < kdim) ? A[abase + k * astride] : (T)0;
      const T b = bcol[(int64_t)k * npad];
      acc = M::run(a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = rt * 16 + M::row(lk, j);
      const int col = ct * 16 + lr;
      if (r < rows && col < ncols) {
        if (x.cplx)
          Cm[r * ld * 2 + col] = acc[j];
        else
          Cm[((r >> 1) * ld + col) * 2 + (r & 1)] = acc[j];
      }
    }
  }
}

// ---------------------------------------------------------------- lane bookkeeping
struct LaneTab {
  int64_t base[MAX_TL];
  double w[MAX_TL];
  int32_t idx[MAX_TL][MAXD];
  int32_t valid[MAX_TL];
};

__device__ void setup_lanes(const PassDesc* __restrict__ d, LaneTab& lt, int64_t tile, int tid) {
  const int tl = d->tl;
  if (tid < tl) {
    const int l = tid;
    const int axis = d->axis;
    const int64_t n = d->n, inner = d->inner;
    int64_t o, i;
    int valid;
    if (inner > 1) {
      o = tile / d->tiles_per_outer;
      i = (tile - o * d->tiles_per_outer) * tl + l;
      valid = i < inner;
    } else {
      o = tile * tl + l;
      i = 0;
      valid = o < d->outer;
    }
    if (!valid) {
      o = 0;
      i = 0;
    }
    lt.valid[l] = valid;
    lt.base[l] = o * n * inner + i;
    // coordinates of the other axes
    int64_t rem = o;
    for (int b = axis - 1; b >= 0; --b) {
      lt.idx[l][b] = (int32_t)(rem % d->shape[b]);
      rem /= d->shape[b];
    }
    rem = i;
    for (int b = d->ndim - 1; b > axis; --b) {
      lt.idx[l][b] = (int32_t)(rem % d->shape[b]);
      rem /= d->shape[b];
    }
    lt.idx[l][axis] = 0;
    double w = 1.0;
    if (d->has_norm) {
      for (int b = 0; b < d->ndim; ++b)
        if (b != axis) w *= d->w_other[b][lt.idx[l][b]];
    }
    lt.w[l] = w;
  }
}

template <typename T, typename T2>
__device__ __forceinline__ void apply_stage(const Stage& st, T2& z, int l, int k, const LaneTab& lt,
                                            const PassDesc* __restrict__ d, double& a0, double& a1, double& a2,
                                            double& a3) {
  const int nops = st.n;
  for (int i = 0; i < nops; ++i) {
    const ElemOp& e = st.op[i];
    switch (e.kind) {
      case EK_MUL1D: {
        const int idx = (e.axis == d->axis) ? k : lt.idx[l][e.axis];
        z = cmul(z, ((const T2*)e.table)[idx]);
        break;
      }
      case EK_MULND:
        z = cmul(z, ((const T2*)e.table)[lt.base[l] + (int64_t)k * d->inner]);
        break;
      case EK_MULB:
        z = cmul(z, ((const T2*)e.table)[0]);
        break;
      case EK_SCALE: {
        const T s = (T)d->S[e.slot];
        z.x *= s;
        z.y *= s;
        break;
      }
      case EK_NORM: {
        const double v = ((double)z.x * (double)z.x + (double)z.y * (double)z.y) * (e.w_axis[k] * lt.w[l]);
        switch (e.acc) {
          case 0: a0 += v; break;
          case 1: a1 += v; break;
          case 2: a2 += v; break;
          default: a3 += v; break;
        }
        break;
      }
      default: break;
    }
  }
}

__device__ void run_scalar_program(const PassDesc* __restrict__ d, const double* acc) {
  double* S = d->S;
  const int ns = d->nscalar;
  for (int i = 0; i < ns; ++i) {
    const ScalarOp& s = d->sc[i];
    switch (s.kind) {
      case SK_SET_ACC:
        S[s.a] = acc[s.b];
        if (!isfinite(acc[s.b]) && d->nonfinite_slot >= 0) S[d->nonfinite_slot] = 1.0;
        break;
      case SK_COPY: S[s.a] = S[s.b]; break;
      case SK_LOSS: {
        S[s.a] = S[s.a] + (S[s.b] - S[s.c]);
        if (s.hist) {
          const double cnt = S[s.cnt_slot];
          const int64_t c = (int64_t)cnt;
          if (c % s.freq == 0 && c / s.freq < s.frames) s.hist[c / s.freq] = S[s.a];
          S[s.cnt_slot] = cnt + 1.0;
        }
        break;
      }
      case SK_SCALE: S[s.a] = sqrt(S[s.b] / S[s.c]); break;
      case SK_CONST: S[s.a] = s.v; break;
      default: break;
    }
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block reduction of the four accumulators, per-tile partials, and the
// last-workgroup finalisation (fixed summation order over tiles).
__device__ void finish_pass(const PassDesc* __restrict__ d, double a0, double a1, double a2, double a3, int tid) {
  __shared__ double red[NTHREADS / 64][MAX_ACC];
  __shared__ int is_last;
  const int64_t tile = blockIdx.x;
  if (!d->has_norm) {
    if (d->nscalar > 0 && tile == 0 && tid == 0) {
      double acc[MAX_ACC] = {0, 0, 0, 0};
      run_scalar_program(d, acc);
    }
    return;
  }
  const int wave = tid >> 6, lane = tid & 63;
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  a3 = wave_sum(a3);
  if (lane == 0) {
    red[wave][0] = a0;
    red[wave][1] = a1;
    red[wave][2] = a2;
    red[wave][3] = a3;
  }
  __syncthreads();
  if (tid == 0) {
    for (int j = 0; j < MAX_ACC; ++j) {
      double s = 0.0;
      for (int w = 0; w < NTHREADS / 64; ++w) s += red[w][j];
      d->partials[tile * MAX_ACC + j] = s;
    }
    __threadfence();
    const unsigned prev = atomicAdd(d->counter, 1u);
    is_last = (prev == (unsigned)(gridDim.x - 1));
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  // fixed-order sum over tiles: strided per thread, then wave and block trees
  const int64_t nt = d->ntiles;
  double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
  for (int64_t t = tid; t < nt; t += NTHREADS) {
    const double* p = d->partials + t * MAX_ACC;
    t0 += __builtin_nontemporal_load(p + 0);
    t1 += __builtin_nontemporal_load(p + 1);
    t2 += __builtin_nontemporal_load(p + 2);
    t3 += __builtin_nontemporal_load(p + 3);
  }
  t0 = wave_sum(t0);
  t1 = wave_sum(t1);
  t2 = wave_sum(t2);
  t3 = wave_sum(t3);
  __syncthreads();
  if (lane == 0) {
    red[wave][0] = t0;
    red[wave][1] = t1;
    red[wave][2] = t2;
    red[wave][3] = t3;
  }
  __syncthreads();
  if (tid == 0) {
    double acc[MAX_ACC];
    for (int j = 0; j < MAX_ACC; ++j) {
      double s = 0.0;
      for (int w = 0; w < NTHREADS / 64; ++w) s += red[w][j];
      acc[j] = s;
    }
    run_scalar_program(d, acc);
    *d->counter = 0u;
    __threadfence();
  }
}

// ---------------------------------------------------------------- the pass kernels
template <typename T>
__global__ __launch_bounds__(NTHREADS) void elem_pass_kernel(const PassDesc* __restrict__ d) {
  using T2 = typename CT<T>::T2;
  __shared__ LaneTab lt;
  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  setup_lanes(d, lt, tile, tid);
  __syncthreads();
  T2* psi = (T2*)d->psi;
  const int tl = d->tl;
  const int n = (int)d->n;
  const int64_t inner = d->inner;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  const int tot = tl * n;
  for (int f = tid; f < tot; f += NTHREADS) {
    int l, k;
    if (inner > 1) {
      l = f % tl;
      k = f / tl;
    } else {
      l = f / n;
      k = f - l * n;
    }
    if (!lt.valid[l]) continue;
    const int64_t addr = lt.base[l] + (int64_t)k * inner;
    T2 z = psi[addr];
    apply_stage<T>(d->st[0], z, l, k, lt, d, a0, a1, a2, a3);
    psi[addr] = z;
  }
  finish_pass(d, a0, a1, a2, a3, tid);
}

template <typename T, int KIND, typename T2>
__device__ __forceinline__ void run_xform(const XformDesc& x, T2*& cur, T2*& oth, int tl, int n, int ld, int tid) {
  if constexpr (KIND == XF_FFT) {
    fft_lds<T>(cur, tl, n, ld, x, tid);
  } else {
    mat_lds<T>(cur, oth, tl, ld, x, tid);
    __syncthreads();
    T2* t = cur;
    cur = oth;
    oth = t;
  }
}

// KIND = XF_FFT: every transform of the pass is an FFT; XF_MAT: every one is a
// matrix product (passes never mix the two).
template <typename T, int KIND>
__global__ __launch_bounds__(NTHREADS) void xform_pass_kernel(const PassDesc* __restrict__ d) {
  using T2 = typename CT<T>::T2;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ LaneTab lt;
  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  setup_lanes(d, lt, tile, tid);
  __syncthreads();
  T2* psi = (T2*)d->psi;
  const int tl = d->tl;
  const int n = (int)d->n;
  const int ld = d->ld;
  const int64_t inner = d->inner;
  const int tot = tl * n;
  T2* cur = (T2*)smem;
  T2* oth = cur + tl * ld;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  // load + stage 0
  for (int f = tid; f < tot; f += NTHREADS) {
    int l, k;
    if (inner > 1) {
      l = f % tl;
      k = f / tl;
    } else {
      l = f / n;
      k = f - l * n;
    }
    T2 z;
    if (lt.valid[l]) {
      z = psi[lt.base[l] + (int64_t)k * inner];
      apply_stage<T>(d->st[0], z, l, k, lt, d, a0, a1, a2, a3);
    } else {
      z.x = 0;
      z.y = 0;
    }
    cur[l * ld + k] = z;
  }
  __syncthreads();
  run_xform<T, KIND>(d->xf[0], cur, oth, tl, n, ld, tid);
  const int nx = d->nx;
  if (nx == 2) {
    if (d->st[1].n > 0) {
      for (int f = tid; f < tot; f += NTHREADS) {
        const int l = f / n, k = f - l * n;
        if (!lt.valid[l]) continue;
        T2 z = cur[l * ld + k];
        apply_stage<T>(d->st[1], z, l, k, lt, d, a0, a1, a2, a3);
        cur[l * ld + k] = z;
      }
      __syncthreads();
    }
    run_xform<T, KIND>(d->xf[1], cur, oth, tl, n, ld, tid);
  }
  // stage nx + store
  const Stage& last = d->st[nx];
  for (int f = tid; f < tot; f += NTHREADS) {
    int l, k;
    if (inner > 1) {
      l = f % tl;
      k = f / tl;
    } else {
      l = f / n;
      k = f - l * n;
    }
    if (!lt.valid[l]) continue;
    T2 z = cur[l * ld + k];
    apply_stage<T>(last, z, l, k, lt, d, a0, a1, a2, a3);
    psi[lt.base[l] + (int64_t)k * inner] = z;
  }
  finish_pass(d, a0, a1, a2, a3, tid);
}

__global__ void scalar_kernel(const PassDesc* __restrict__ d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double acc[MAX_ACC] = {0, 0, 0, 0};
    run_scalar_program(d, acc);
  }
}

// ---------------------------------------------------------------- observers
// state_density (wave_function.rs:121-148): for ndim > 1 the marginal along
// `axis` of |psi|^2 prod_all w; one workgroup per index, fixed-order reduction.
template <typename T>
__global__ __launch_bounds__(NTHREADS) void state_density_kernel(const void* psi_, int ndim, int64_t s0, int64_t s1,
                                                                 int64_t s2, int64_t s3, int axis, const double* w0,
                                                                 const double* w1, const double* w2,
                                                                 const double* w3, double* out, int64_t ostride) {
  using T2 = typename CT<T>::T2;
  __shared__ double red[NTHREADS];
  const T2* psi = (const T2*)psi_;
  const int64_t shape[MAXD] = {s0, s1, s2, s3};
  const double* w[MAXD] = {w0, w1, w2, w3};
  const int64_t k = blockIdx.x;
  int64_t outer = 1, inner = 1;
  for (int b = 0; b < axis; ++b) outer *= shape[b];
  for (int b = axis + 1; b < ndim; ++b) inner *= shape[b];
  const int64_t n = shape[axis];
  double acc = 0.0;
  const int64_t m_tot = outer * inner;
  for (int64_t m = threadIdx.x; m < m_tot; m += NTHREADS) {
    const int64_t o = m / inner, i = m - o * inner;
    double wt = w[axis][k];
    int64_t rem = o;
    for (int b = axis - 1; b >= 0; --b) {
      wt *= w[b][rem % shape[b]];
      rem /= shape[b];
    }
    rem = i;
    for (int b = ndim - 1; b > axis; --b) {
      wt *= w[b][rem % shape[b]];
      rem /= shape[b];
    }
    const T2 z = psi[(o * n + k) * inner + i];
    acc += ((double)z.x * (double)z.x + (double)z.y * (double)z.y) * wt;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = NTHREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k * ostride] = red[0];
}

// density (wave_function.rs:112-118): unweighted |psi|^2
template <typename T>
__global__ void density_kernel(const void* psi_, int64_t size, double* out, int64_t ostride) {
  using T2 = typename CT<T>::T2;
  const T2* psi = (const T2*)psi_;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < size; i += (int64_t)gridDim.x * blockDim.x) {
    const T2 z = psi[i];
    out[i * ostride] = (double)z.x * (double)z.x + (double)z.y * (double)z.y;
  }
}

// dot (wave_function.rs:87-101): sum a conj(b) prod w, FP64, deterministic
template <typename T>
__global__ __launch_bounds__(NTHREADS) void dot_kernel(const void* a_, const void* b_, int ndim, int64_t s0, int64_t s1,
                                                       int64_t s2, int64_t s3, const double* w0, const double* w1,
                                                       const double* w2, const double* w3, double* partials,
                                                       unsigned int* counter, double* out) {
  using T2 = typename CT<T>::T2;
  __shared__ double rr[NTHREADS], ri[NTHREADS];
  __shared__ int is_last;
  const T2* a = (const T2*)a_;
  const T2* b = (const T2*)b_;
  const int64_t shape[MAXD] = {s0, s1, s2, s3};
  const double* w[MAXD] = {w0, w1, w2, w3};
  int64_t size = 1;
  for (int q = 0; q < ndim; ++q) size *= shape[q];
  double accr = 0.0, acci = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)NTHREADS + threadIdx.x; i < size; i += (int64_t)gridDim.x * NTHREADS) {
    double wt = 1.0;
    int64_t rem = i;
    for (int q = ndim - 1; q >= 0; --q) {
      wt *= w[q][rem % shape[q]];
      rem /= shape[q];
    }
    const T2 x = a[i], y = b[i];
    accr += ((double)x.x * (double)y.x + (double)x.y * (double)y.y) * wt;
    acci += ((double)x.y * (double)y.x - (double)x.x * (double)y.y) * wt;
  }
  rr[threadIdx.x] = accr;
  ri[threadIdx.x] = acci;
  __syncthreads();
  for (int st = NTHREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      rr[threadIdx.x] += rr[threadIdx.x + st];
      ri[threadIdx.x] += ri[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = rr[0];
    partials[2 * blockIdx.x + 1] = ri[0];
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (is_last && threadIdx.x == 0) {
    __threadfence();
    double sr = 0.0, si = 0.0;
    for (unsigned q = 0; q < gridDim.x; ++q) {
      sr += __builtin_nontemporal_load(&partials[2 * q]);
      si += __builtin_nontemporal_load(&partials[2 * q + 1]);
    }
    out[0] = sr;
    out[1] = si;
    *counter = 0u;
    __threadfence();
  }
}

// f64 complex <-> T complex
template <typename T>
__global__ void convert_kernel(const double* src, T* dst, int64_t n2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (T)src[i];
}
template <typename T>
__global__ void convert_back_kernel(const T* src, double* dst, int64_t n2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (double)src[i];
}

// Gauss-Legendre nodes (ascending) and weights in FP64: Newton on P_n, one
// thread per root. Not in the crate (its callers supply the polar grid).
__global__ void gauss_legendre_kernel(int64_t n, double* x, double* w) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double z = cos(M_PI * ((double)i + 0.75) / ((double)n + 0.5));
  double p0 = 1.0, p1 = z, dp = 1.0;
  for (int it = 0; it < 100; ++it) {
    p0 = 1.0;
    p1 = z;
    for (int64_t k = 2; k <= n; ++k) {
      const double p2 = ((double)(2 * k - 1) * z * p1 - (double)(k - 1) * p0) / (double)k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) {
      p0 = 1.0;
      p1 = z;
    }
    dp = (double)n * (z * p1 - p0) / (z * z - 1.0);
    const double dz = p1 / dp;
    z -= dz;
    if (fabs(dz) < 1e-16) break;
  }
  p0 = 1.0;
  p1 = z;
  for (int64_t k = 2; k <= n; ++k) {
    const double p2 = ((double)(2 * k - 1) * z * p1 - (double)(k - 1) * p0) / (double)k;
    p0 = p1;
    p1 = p2;
  }
  if (n == 1) p0 = 1.0;
  dp = (double)n * (z * p1 - p0) / (z * z - 1.0);
  x[n - 1 - i] = z;
  w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
}

}  // namespace

// ---------------------------------------------------------------- host launchers
static size_t xform_lds_bytes(int dtype, const PassDesc& h) {
  const size_t csz = dtype ? 16 : 8;
  size_t bufs = 1;
  for (int i = 0; i < h.nx; ++i)
    if (h.xf[i].type == XF_MAT) bufs = 2;
  return bufs * (size_t)h.tl * (size_t)h.ld * csz;
}

int sod_launch_pass(int dtype, const PassDesc* d_desc, const PassDesc& h, hipStream_t s) {
  const dim3 grid((unsigned)h.ntiles), block(NTHREADS);
  if (h.nx == 0) {
    if (dtype)
      elem_pass_kernel<double><<<grid, block, 0, s>>>(d_desc);
    else
      elem_pass_kernel<float><<<grid, block, 0, s>>>(d_desc);
  } else {
    const size_t lds = xform_lds_bytes(dtype, h);
    static bool attr_set = false;
    if (!attr_set) {
      (void)hipFuncSetAttribute((const void*)xform_pass_kernel<double, XF_FFT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)SOD_MAX_DYN_LDS);
      (void)hipFuncSetAttribute((const void*)xform_pass_kernel<float, XF_FFT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)SOD_MAX_DYN_LDS);
      (void)hipFuncSetAttribute((const void*)xform_pass_kernel<double, XF_MAT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)SOD_MAX_DYN_LDS);
      (void)hipFuncSetAttribute((const void*)xform_pass_kernel<float, XF_MAT>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)SOD_MAX_DYN_LDS);
      attr_set = true;
    }
    if (lds > SOD_MAX_DYN_LDS) return (int)hipErrorInvalidValue;
    const bool mat = h.xf[0].type == XF_MAT;
    if (dtype) {
      if (mat)
        xform_pass_kernel<double, XF_MAT><<<grid, block, lds, s>>>(d_desc);
      else
        xform_pass_kernel<double, XF_FFT><<<grid, block, lds, s>>>(d_desc);
    } else {
      if (mat)
        xform_pass_kernel<float, XF_MAT><<<grid, block, lds, s>>>(d_desc);
      else
        xform_pass_kernel<float, XF_FFT><<<grid, block, lds, s>>>(d_desc);
    }
  }
  return (int)hipGetLastError();
}

int sod_launch_scalar(const PassDesc* d_desc, hipStream_t s) {
  scalar_kernel<<<1, 64, 0, s>>>(d_desc);
  return (int)hipGetLastError();
}

int sod_launch_state_density(int dtype, const void* psi, int ndim, const int64_t* shape, int axis,
                             const double* const* w, double* out, hipStream_t s) {
  int64_t sh[MAXD] = {1, 1, 1, 1};
  const double* ww[MAXD] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < ndim; ++i) {
    sh[i] = shape[i];
    ww[i] = w[i];
  }
  const dim3 grid((unsigned)sh[axis]);
  if (dtype)
    state_density_kernel<double><<<grid, NTHREADS, 0, s>>>(psi, ndim, sh[0], sh[1], sh[2], sh[3], axis, ww[0], ww[1],
                                                           ww[2], ww[3], out, (int64_t)1);
  else
    state_density_kernel<float><<<grid, NTHREADS, 0, s>>>(psi, ndim, sh[0], sh[1], sh[2], sh[3], axis, ww[0], ww[1],
                                                          ww[2], ww[3], out, (int64_t)1);
  return (int)hipGetLastError();
}

int sod_launch_density(int dtype, const void* psi, int64_t size, double* out, int64_t ostride, hipStream_t s) {
  int64_t blocks = (size + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  if (dtype)
    density_kernel<double><<<(unsigned)blocks, 256, 0, s>>>(psi, size, out, ostride);
  else
    density_kernel<float><<<(unsigned)blocks, 256, 0, s>>>(psi, size, out, ostride);
  return (int)hipGetLastError();
}

int sod_launch_dot(int dtype, const void* a, const void* b, int ndim, const int64_t* shape, const double* const* w,
                   double* partials, unsigned int* counter, double* out, hipStream_t s) {
  int64_t sh[MAXD] = {1, 1, 1, 1};
  const double* ww[MAXD] = {nullptr, nullptr, nullptr, nullptr};
  int64_t size = 1;
  for (int i = 0; i < ndim; ++i) {
    sh[i] = shape[i];
    ww[i] = w[i];
    size *= shape[i];
  }
  int64_t blocks = (size + NTHREADS - 1) / NTHREADS;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  if (dtype)
    dot_kernel<double><<<(unsigned)blocks, NTHREADS, 0, s>>>(a, b, ndim, sh[0], sh[1], sh[2], sh[3], ww[0], ww[1],
                                                             ww[2], ww[3], partials, counter, out);
  else
    dot_kernel<float><<<(unsigned)blocks, NTHREADS, 0, s>>>(a, b, ndim, sh[0], sh[1], sh[2], sh[3], ww[0], ww[1],
                                                            ww[2], ww[3], partials, counter, out);
  return (int)hipGetLastError();
}

int sod_launch_convert(int dtype, const void* src, void* dst, int64_t n_complex, int direction, hipStream_t s) {
  const int64_t n2 = 2 * n_complex;
  int64_t blocks = (n2 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  if (direction == 0) {  // f64 -> T
    if (dtype)
      convert_kernel<double><<<(unsigned)blocks, 256, 0, s>>>((const double*)src, (double*)dst, n2);
    else
      convert_kernel<float><<<(unsigned)blocks, 256, 0, s>>>((const double*)src, (float*)dst, n2);
  } else {  // T -> f64
    if (dtype)
      convert_back_kernel<double><<<(unsigned)blocks, 256, 0, s>>>((const double*)src, (double*)dst, n2);
    else
      convert_back_kernel<float><<<(unsigned)blocks, 256, 0, s>>>((const float*)src, (double*)dst, n2);
  }
  return (int)hipGetLastError();
}

int sod_launch_gauss_legendre(int64_t n, double* x, double* w, hipStream_t s) {
  const unsigned blocks = (unsigned)((n + 127) / 128);
  gauss_legendre_kernel<<<blocks, 128, 0, s>>>(n, x, w);
  return (int)hipGetLastError();
}
