// Element contraction on the matrix cores: E = Psi_t^T T with T = D^ Psi_c.
//
// One wavefront contracts one element.  The reduction index K runs over (quadrature point
// q, active test row r): K = q * nrt + (r - rt0), with the active rows of the block (the
// value row, the three reference-gradient rows, or all four; the reference skips the
// identically zero pointwise blocks the same way, kernel.py:1606-1646).  K is padded to a
// multiple of 4 and the element's nd x nd output to 16 x 16 tiles of
// v_mfma_f64_16x16x4f64 (FP64) or v_mfma_f32_16x16x4f32 (FP32):
//   A[i][k] (16 x 4): lane l holds A[l % 16][l / 16]   -- Psi_t, constant, in LDS
//   B[k][j] (4 x 16): lane l holds B[l / 16][l % 16]   -- T, per element, in LDS
//   D[i][j] (16x16): lane l, register v holds D[row(l / 16, v)][l % 16], with
//     row(g, v) = g + 4 v for the FP64 instruction and 4 g + v for the FP32 one.
// The geometry of the element map (J, det J, J^-1), the fields and the pointwise blocks D^
// are FP64 whatever the precision of the contraction; T is rounded to it.
#include <hip/hip_runtime.h>

#include "hm_common.h"
#include "hm_device.h"
#include "hm_point.h"

namespace hm {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename VT>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4v acc;
  __device__ static inline acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  __device__ static inline acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int g, int v) { return g + 4 * v; }
};
template <>
struct Mfma<float> {
  typedef f4v acc;
  __device__ static inline acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  __device__ static inline acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int g, int v) { return 4 * g + v; }
};

// Shape of one block's contraction.
struct MfmaShape {
  int nq, nd_t, nd_c;
  int rt0, nrt, ct0, nct;  // active test rows [rt0, rt0+nrt), trial columns [ct0, ct0+nct)
  int Kv, Kp;              // valid and padded reduction length
  int wpb;                 // wavefronts (elements) per block
};

__device__ inline double psi_value(const double* tab, int nq, int nd, int q, int r, int a) {
  return r == 0 ? tab[q * nd + a] : tab[(size_t)nq * nd + ((size_t)q * nd + a) * 3 + (r - 1)];
}

template <typename VT, int MT, int NT>
__global__ void __launch_bounds__(256) k_element_mfma(hm_problem P, int s, int j, int64_t e0, int64_t e1,
                                                      MfmaShape S, void* out) {
  extern __shared__ double lds[];
  constexpr int MW = MT * 16, NW = NT * 16;
  const int nq = S.nq, nd_t = S.nd_t, nd_c = S.nd_c, Kp = S.Kp, Kv = S.Kv;
  const double* tab_t = s == HM_PARAMETER ? P.tab_m : P.tab_u;
  const double* tab_c = j == HM_PARAMETER ? P.tab_m : P.tab_u;
  // LDS: A table (VT, Kp x MW), trial table (FP64, nq x nct x NW), then per wave
  // D (FP64, nq x 16), T (VT, Kp x NW), element data.
  VT* At = (VT*)lds;
  double* Pc = (double*)(At + (size_t)Kp * MW + ((Kp * MW) & 1));
  double* wave_base = Pc + (size_t)nq * S.nct * NW;
  const size_t elem_words = 24 + 2 * P.nd_u + P.nd_m;
  const size_t t_words = ((size_t)Kp * NW * sizeof(VT) + 7) / 8;
  const size_t wave_words = (size_t)nq * 16 + t_words + elem_words;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* D = wave_base + wave * wave_words;
  VT* T = (VT*)(D + (size_t)nq * 16);
  double* Xe = D + (size_t)nq * 16 + t_words;
  double* ue = Xe + 24;
  double* pe = ue + P.nd_u;
  double* me = pe + P.nd_u;

  for (int i = threadIdx.x; i < Kp * MW; i += blockDim.x) {
    const int K = i / MW, a = i % MW;
    double v = 0.0;
    if (K < Kv && a < nd_t) v = psi_value(tab_t, nq, nd_t, K / S.nrt, S.rt0 + K % S.nrt, a);
    At[i] = (VT)v;
  }
  for (int i = threadIdx.x; i < nq * S.nct * NW; i += blockDim.x) {
    const int b = i % NW, ci = (i / NW) % S.nct, q = i / (NW * S.nct);
    Pc[i] = b < nd_c ? psi_value(tab_c, nq, nd_c, q, S.ct0 + ci, b) : 0.0;
  }

  for (int64_t base = e0 + (int64_t)blockIdx.x * S.wpb; base < e1; base += (int64_t)gridDim.x * S.wpb) {
    const int64_t e = base + wave;
    const bool active = wave < S.wpb && e < e1;
    __syncthreads();
    if (active) {
      for (int i = lane; i < 24; i += 64) Xe[i] = P.verts[(size_t)P.elem_verts[e * 8 + i / 3] * 3 + i % 3];
      for (int a = lane; a < P.nd_u; a += 64) {
        const int32_t d = P.edofs_u[e * P.nd_u + a];
        ue[a] = P.u[d];
        pe[a] = P.p[d];
      }
      for (int a = lane; a < P.nd_m; a += 64) me[a] = P.m[P.edofs_m[e * P.nd_m + a]];
    }
    __syncthreads();
    if (active) {
      for (int q = lane; q < nq; q += 64) {
        // (the FP32 form: fields, kappa and the pairs in FP32, the geometry in FP64)
        PointStateT<VT> st;
        point_state<VT>(P, q, Xe, ue, me, pe, st);
        VT Dq[4][4];
        point_block(s, j, st, Dq);
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) D[q * 16 + r * 4 + c] = Dq[r][c];
      }
    }
    __syncthreads();
    if (active) {
      // T[K][b] = sum_c D[q][r][c] Psi_c[q][c][b]
      for (int i = lane; i < Kp * NW; i += 64) {
        const int K = i / NW, b = i % NW;
        VT v = 0;
        if (K < Kv) {
          const int q = K / S.nrt, r = S.rt0 + K % S.nrt;
          const double* Dr = D + q * 16 + r * 4 + S.ct0;
          const double* pc = Pc + (size_t)q * S.nct * NW + b;
          for (int ci = 0; ci < S.nct; ++ci) v = fma((VT)Dr[ci], (VT)pc[ci * NW], v);
        }
        T[i] = v;
      }
    }
    __syncthreads();
    if (active) {
      const int g = lane >> 4, c16 = lane & 15;
      typename Mfma<VT>::acc acc[MT][NT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = Mfma<VT>::zero();
      for (int K = g; K < Kp; K += 4) {
        VT av[MT], bv[NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) av[mt] = At[K * MW + mt * 16 + c16];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bv[nt] = T[K * NW + nt * 16 + c16];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = Mfma<VT>::op(av[mt], bv[nt], acc[mt][nt]);
      }
      VT* Eo = (VT*)out + (size_t)(e - e0) * nd_t * nd_c;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = mt * 16 + Mfma<VT>::row(g, v), col = nt * 16 + c16;
            if (row < nd_t && col < nd_c) Eo[row * nd_c + col] = acc[mt][nt][v];
          }
    }
  }
}

static size_t shape_bytes(const MfmaShape& S, const hm_problem& P, size_t vt_size) {
  const int MW = ((S.nd_t + 15) / 16) * 16, NW = ((S.nd_c + 15) / 16) * 16;
  const size_t at_words = ((size_t)S.Kp * MW * vt_size + 7) / 8 + 1;
  const size_t pc_words = (size_t)S.nq * S.nct * NW;
  const size_t t_words = ((size_t)S.Kp * NW * vt_size + 7) / 8;
  const size_t wave_words = (size_t)S.nq * 16 + t_words + 24 + 2 * P.nd_u + P.nd_m;
  return (at_words + pc_words + (size_t)S.wpb * wave_words) * 8;
}

static int max_lds_per_block() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 64 * 1024;
  if (cached[dev] == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 64 * 1024;
    cached[dev] = v;
  }
  return cached[dev];
}

template <typename VT, int MT, int NT>
static int launch_mfma(const hm_problem& P, int i, int j, int64_t e0, int64_t e1, MfmaShape S, void* out,
                       hipStream_t st) {
  auto kern = k_element_mfma<VT, MT, NT>;
  const int limit = max_lds_per_block();
  size_t bytes = 0;
  for (S.wpb = 4; S.wpb >= 1; --S.wpb) {
    bytes = shape_bytes(S, P, sizeof(VT));
    if (bytes <= (size_t)limit) break;
  }
  if (S.wpb < 1) HM_FAIL(HM_ERR_UNSUPPORTED, "the contraction tables need %zu bytes of LDS", bytes);
  if (bytes > 64 * 1024) {
    hipError_t err = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (err != hipSuccess) HM_FAIL(HM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(err));
  }
  const int64_t groups = (e1 - e0 + S.wpb - 1) / S.wpb;
  const int grid = (int)(groups < 2048 ? groups : 2048);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), bytes, st, P, i, j, e0, e1, S, out);
  return hip_check("k_element_mfma");
}

}  // namespace hm

using namespace hm;

// Called by hm_element_matrices for nd <= 32; precision HM_FP64 or HM_FP32.
int hm_element_matrices_mfma(const hm_problem* P, int i, int j, int64_t e0, int64_t e1, int precision,
                             void* out, void* stream) {
  MfmaShape S;
  S.nq = P->nq;
  S.nd_t = i == HM_PARAMETER ? P->nd_m : P->nd_u;
  S.nd_c = j == HM_PARAMETER ? P->nd_m : P->nd_u;
  const int parts = block_parts(i, j, P->density);
  if (parts == 0) {
    // an identically zero block (W_uu of the linear density)
    const size_t n = (size_t)(e1 - e0) * S.nd_t * S.nd_c;
    return hm_fill_zero(out, n, precision == HM_FP64 ? 8 : 4, (hipStream_t)stream);
  }
  S.rt0 = (parts & 1) ? 0 : 1;
  S.nrt = ((parts & 2) ? 4 : 1) - S.rt0;
  S.ct0 = (parts & 4) ? 0 : 1;
  S.nct = ((parts & 8) ? 4 : 1) - S.ct0;
  S.Kv = S.nq * S.nrt;
  S.Kp = (S.Kv + 3) / 4 * 4;
  S.wpb = 4;
  const int MT = (S.nd_t + 15) / 16, NT = (S.nd_c + 15) / 16;
  hipStream_t st = (hipStream_t)stream;
  if (precision == HM_FP64) {
    if (MT == 2 && NT == 2) return launch_mfma<double, 2, 2>(*P, i, j, e0, e1, S, out, st);
    if (MT == 2 && NT == 1) return launch_mfma<double, 2, 1>(*P, i, j, e0, e1, S, out, st);
    if (MT == 1 && NT == 2) return launch_mfma<double, 1, 2>(*P, i, j, e0, e1, S, out, st);
    return launch_mfma<double, 1, 1>(*P, i, j, e0, e1, S, out, st);
  }
  if (MT == 2 && NT == 2) return launch_mfma<float, 2, 2>(*P, i, j, e0, e1, S, out, st);
  if (MT == 2 && NT == 1) return launch_mfma<float, 2, 1>(*P, i, j, e0, e1, S, out, st);
  if (MT == 1 && NT == 2) return launch_mfma<float, 1, 2>(*P, i, j, e0, e1, S, out, st);
  return launch_mfma<float, 1, 1>(*P, i, j, e0, e1, S, out, st);
}

namespace hm {
__global__ void k_fill_zero8(uint64_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
__global__ void k_fill_zero4(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
// *flag = 1 where a value is not finite (an exponent of all ones: infinity or NaN)
template <typename T>
__global__ void k_flag_nonfinite(const T* __restrict__ p, size_t n, int32_t* flag) {
  bool bad = false;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    bad |= !__builtin_isfinite(p[i]);
  if (bad) atomicOr(flag, 1);
}
}  // namespace hm

int hm_flag_nonfinite(const void* p, size_t n, int word_bytes, int32_t* flag, hipStream_t st) {
  if (n == 0 || flag == nullptr) return HM_OK;
  const size_t nb = (n + 255) / 256;
  const int grid = (int)(nb < 8192 ? nb : 8192);
  if (word_bytes == 8)
    hipLaunchKernelGGL(k_flag_nonfinite<double>, dim3(grid), dim3(256), 0, st, (const double*)p, n, flag);
  else
    hipLaunchKernelGGL(k_flag_nonfinite<float>, dim3(grid), dim3(256), 0, st, (const float*)p, n, flag);
  return hip_check("flag non-finite");
}

int hm_fill_zero(void* p, size_t n, int word_bytes, hipStream_t st) {
  if (n == 0) return HM_OK;
  const size_t nb = (n + 255) / 256;
  const int grid = (int)(nb < 65536 ? nb : 65536);
  if (word_bytes == 8)
    hipLaunchKernelGGL(k_fill_zero8, dim3(grid), dim3(256), 0, st, (uint64_t*)p, n);
  else
    hipLaunchKernelGGL(k_fill_zero4, dim3(grid), dim3(256), 0, st, (uint32_t*)p, n);
  return hip_check("fill zero");
}
