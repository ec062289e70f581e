// orbital_gemm.hip -- kernel (b): the orbital contraction on MFMA.
//
// Reference: QMCSlater::evaluate's five products Xw * C^T (QMCSlater.cpp:
// 856-862, Array2D::gemm with rhsIsTransposed, Array2D.h:425-504).  All five
// components (value, d/dx, d/dy, d/dz, Laplacian) and every point of a batch
// form one GEMM, computed transposed so both operands stream coalesced:
//
//   D^T[k][o][p] = sum_b  Ct[b][o] * X[k][b][p]
//
// Ct = C^T ([basis][orbital], zero padded to the orbital tile widths and to
// 32-multiples of basis functions; BF16 operands keep C orbital-major,
// [orbital][basis], so a lane's eight k values of one orbital are one 16-byte
// LDS read), X from kernel (a).  One wavefront owns a 16-point tile and MT
// 16-orbital tiles; each k-step issues MT MFMA instructions: 16x16x4 for
// FP64 (v_mfma_f64_16x16x4f64) and FP32 (v_mfma_f32_16x16x4f32) with lane l
// supplying A[l%16][l/16] and B[l/16][l%16]; gfx950's 16x16x32 for BF16
// operands (v_mfma_f32_16x16x32_bf16, FP32 accumulate) with lane l supplying
// the eight k values 8(l/16)..+7 of A[l%16][.] and B[.][l%16] (any k
// assignment shared by A and B gives the same sum).  Accumulator element i
// (i = 0..3) of lane l is D[row][l%16] with row = 4(l/16) + i for the
// FP32-output instructions and row = 4i + l/16 for the FP64 one
// (Mfma<T>::row).
#include <hip/hip_bf16.h>

#include "qmcb_kernels.hpp"

namespace qmcb {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));

template <typename TI>
struct Mfma;

template <>
struct Mfma<double> {
  typedef d4v acc_t;
  typedef double acc_s;
  static constexpr int KPL = 1;  // k values per lane per instruction
  __device__ static acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output row of accumulator element i in lane group lk = lane / 16
  __device__ static constexpr int row(int lk, int i) { return 4 * i + lk; }
};

template <>
struct Mfma<float> {
  typedef f4v acc_t;
  typedef float acc_s;
  static constexpr int KPL = 1;
  __device__ static acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int lk, int i) { return 4 * lk + i; }
};

template <>
struct Mfma<__hip_bfloat16> {
  typedef f4v acc_t;
  typedef float acc_s;
  static constexpr int KPL = 8;
  __device__ static acc_t run(b8v a, b8v b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int lk, int i) { return 4 * lk + i; }
};

// Coefficient k-chunks are staged in LDS once per workgroup (WPB waves, 2 or
// 4, share them) and double-buffered: while the waves run the MFMAs of chunk c from
// LDS, the next chunk is loaded from L2 into registers.  The point operand
// X streams from HBM exactly once.  LDS rows are padded so that the four
// k-rows one MFMA operand read touches (lanes 0-15 / 16-31 / ...) fall in
// different banks; BF16 chunks are stored orbital-major (KC + 8 elements a
// row: the 16 orbitals of one 16-byte read start 20 banks apart and cover
// the 64 banks once).
template <typename TI>
struct GemmTile {
  static constexpr int KC = 32;                       // k rows per chunk
  static constexpr int PAD = sizeof(TI) == 8 ? 8 : 16;  // elements
};

template <typename TI, typename TO, int MT, int WPB>
__global__ void __launch_bounds__(64 * WPB)
orbital_gemm_kernel(const TI *__restrict__ Ct, int ldc, int n_orb, const TI *__restrict__ X,
                    int ldk, long ldp, long P, int kpad, GemmOut o) {
  typedef Mfma<TI> M;
  typedef typename M::acc_t acc_t;
  constexpr int KC = GemmTile<TI>::KC;
  constexpr int OW = 16 * MT;                    // orbital columns of the tile
  constexpr bool KT = M::KPL > 1;                // chunk stored orbital-major
  constexpr int LR = KT ? OW : KC;               // LDS rows
  constexpr int LW = KT ? KC + 8 : OW + GemmTile<TI>::PAD;  // LDS row length
  constexpr int VE = 16 / sizeof(TI);            // elements per 16-byte load
  constexpr int NT = 64 * WPB;
  constexpr int NV = (KC * OW + NT * VE - 1) / (NT * VE);  // 16-byte loads per thread, chunk
  static_assert(KC * OW % VE == 0, "tile / vector size mismatch");
  __shared__ TI cs[2][LR][LW];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const long p0 = ((long)blockIdx.x * WPB + wave) * 16;
  const bool live = p0 < P;  // wave-uniform
  const int comp = blockIdx.y;
  const int o0 = blockIdx.z * OW;
  const int li = lane & 15, lk = lane >> 4;
  acc_t acc[MT];
#pragma unroll
  for (int t = 0; t < MT; t++) acc[t] = acc_t{0, 0, 0, 0};

  // cooperative chunk load: element e = (j * NT + tid) * VE of the KC x OW
  // tile (OW x KC for orbital-major BF16, whose rows are kpad long)
  typedef __attribute__((ext_vector_type(4))) unsigned int u4;
  u4 stage[NV];
  auto gload = [&](int k0) {
#pragma unroll
    for (int j = 0; j < NV; j++) {
      const int e = (j * NT + tid) * VE, r = e / (LR == KC ? OW : KC), c = e % (LR == KC ? OW : KC);
      const long src = KT ? (long)(o0 + r) * kpad + k0 + c : (long)(k0 + r) * ldc + o0 + c;
      if (e < KC * OW) stage[j] = *(const u4 *)(Ct + src);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < NV; j++) {
      const int e = (j * NT + tid) * VE, r = e / (LR == KC ? OW : KC), c = e % (LR == KC ? OW : KC);
      if (e < KC * OW) *(u4 *)&cs[buf][r][c] = stage[j];
    }
  };
  constexpr int KPL = M::KPL, KS = 4 * KPL;  // k values per lane, per instruction
  const TI *__restrict__ xp = X + (long)comp * ldk * ldp + (long)(KPL * lk) * ldp + p0 + li;
  const int nch = kpad / KC;
  // the point operand of the next chunk is loaded with the next coefficient
  // chunk, one chunk ahead of its MFMAs (its HBM latency overlaps them)
  TI b[KC / 4], bn[KC / 4];
  auto bload = [&](int k0, TI *dst) {
    if (!live) return;
#pragma unroll
    for (int s = 0; s < KC / KS; s++)
#pragma unroll
      for (int j = 0; j < KPL; j++) dst[s * KPL + j] = xp[(long)(k0 + KS * s + j) * ldp];
  };
  gload(0);
  bload(0, b);
  lstore(0);
  __syncthreads();
  for (int ch = 0; ch < nch; ch++) {
    const int buf = ch & 1, k0 = ch * KC;
    if (ch + 1 < nch) {
      gload(k0 + KC);
      bload(k0 + KC, bn);
    }
    if (live) {
      if constexpr (KPL == 1) {
#pragma unroll
        for (int s = 0; s < KC / 4; s++) {
          TI a[MT];
#pragma unroll
          for (int t = 0; t < MT; t++) a[t] = cs[buf][4 * s + lk][16 * t + li];
#pragma unroll
          for (int t = 0; t < MT; t++) acc[t] = M::run(a[t], b[s], acc[t]);
        }
      } else {
#pragma unroll
        for (int s = 0; s < KC / KS; s++) {
          b8v bv;
#pragma unroll
          for (int j = 0; j < KPL; j++) bv[j] = __builtin_bit_cast(__bf16, b[s * KPL + j]);
#pragma unroll
          for (int t = 0; t < MT; t++) {
            const b8v av = *(const b8v *)&cs[buf][16 * t + li][KS * s + KPL * lk];
            acc[t] = M::run(av, bv, acc[t]);
          }
        }
      }
    }
    if (ch + 1 < nch) {
      lstore(buf ^ 1);
#pragma unroll
      for (int s = 0; s < KC / 4; s++) b[s] = bn[s];
    }
    __syncthreads();
  }
  TO *out = (TO *)o.out;
  if (o.s_orb == 1 && o.per_walker == 1 && !o.walker_map) {
    // point-major output (one walker's orbital row contiguous): transpose each
    // 16x16 tile through LDS so a point's 16 orbitals go out as one segment
    __shared__ TO tile[WPB][16][17];
#pragma unroll
    for (int t = 0; t < MT; t++) {
#pragma unroll
      for (int i = 0; i < 4; i++) tile[wave][li][M::row(lk, i)] = (TO)acc[t][i];
      __syncthreads();
      const int pp = lane >> 2, oo = (lane & 3) * 4;
      const long p = p0 + pp;
      if (live && p < P) {
        TO *dst = out + p * o.s_walker + (long)comp * o.s_comp + o0 + 16 * t + oo;
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (o0 + 16 * t + oo + q < n_orb) dst[q] = tile[wave][pp][oo + q];
      }
      __syncthreads();
    }
    return;
  }
  const long p = p0 + li;
  if (!live || p >= P) return;
  const long row = p / o.per_walker;
  const long wk = o.walker_map ? o.walker_map[row] : row;
  const long pe = (p % o.per_walker) * o.s_elec;
  if (o.out_val && comp == 0) {
    double *ov = o.out_val + wk * o.s_walker_val + (p % o.per_walker) * o.s_elec_val;
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int orb = o0 + 16 * t + M::row(lk, i);
        if (orb < n_orb) ov[(long)orb * o.s_orb_val] = (double)acc[t][i];
      }
    return;
  }
  const int oc = o.out_val ? comp - 1 : comp;
  const long base = (long)oc * o.s_comp + wk * o.s_walker + pe;
#pragma unroll
  for (int t = 0; t < MT; t++) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int orb = o0 + 16 * t + M::row(lk, i);
      if (orb < n_orb) out[base + (long)orb * o.s_orb] = (TO)acc[t][i];
    }
  }
}

// Orbital tiles.  Batches of up to kSplitRows points (single-electron moves:
// one row per walker) use tiles of 64 orbital columns, wider orbital sets
// split over grid.z, and each workgroup re-reads its points' X rows, mostly
// from L2.  At 128 orbitals and 8192 rows two 64-column tiles beat one
// 128-column tile even alone (FP32 155 against 162 us, FP64 260 against
// 316 us): twice the workgroups at 76 / 138 registers and 20 / 37 KB of LDS
// each, against 130 / 220 and 41 / 78 KB, and their waves fit beside
// sm_move's.  Larger batches (the all-electron evaluate, ~10^5 rows per
// launch) stream X from HBM, where reading it once with 128-column tiles is
// faster (1418 against 1702 us FP32 per 88320-row chunk).
constexpr long kSplitRows = 16384;
int gemm_orbital_tiles(int n_orb, long P) {
  if (n_orb > 64 && P > kSplitRows) return 8;
  if (n_orb > 32) return 4;
  if (n_orb > 16) return 2;
  return 1;
}

bool gemm_orbital_tiles_valid(int mt, int n_orb, int ldc) {
  if (mt != 1 && mt != 2 && mt != 4 && mt != 8) return false;
  const int ow = 16 * mt;
  return (long)((n_orb + ow - 1) / ow) * ow <= ldc;  // every tile inside the padded orbitals
}

template <typename TI, typename TO, int WPB>
static void gemm_grid(const TI *Ct, int ldc, int n_orb, const TI *X, int ldk, long ldp, long P,
                      int kpad, const GemmOut &o, hipStream_t st, int mt) {
  const unsigned gx = (unsigned)((P + 16 * WPB - 1) / (16 * WPB));
  const dim3 blk(64 * WPB);
  if (mt == 8) {
    hipLaunchKernelGGL((orbital_gemm_kernel<TI, TO, 8, WPB>),
                       dim3(gx, kComp, (unsigned)((n_orb + 127) / 128)), blk, 0, st, Ct, ldc,
                       n_orb, X, ldk, ldp, P, kpad, o);
  } else if (mt == 4) {
    hipLaunchKernelGGL((orbital_gemm_kernel<TI, TO, 4, WPB>),
                       dim3(gx, kComp, (unsigned)((n_orb + 63) / 64)), blk, 0, st, Ct, ldc, n_orb,
                       X, ldk, ldp, P, kpad, o);
  } else if (mt == 2) {
    hipLaunchKernelGGL((orbital_gemm_kernel<TI, TO, 2, WPB>),
                       dim3(gx, kComp, (unsigned)((n_orb + 31) / 32)), blk, 0, st, Ct, ldc, n_orb,
                       X, ldk, ldp, P, kpad, o);
  } else {
    hipLaunchKernelGGL((orbital_gemm_kernel<TI, TO, 1, WPB>),
                       dim3(gx, kComp, (unsigned)((n_orb + 15) / 16)), blk, 0, st, Ct, ldc, n_orb,
                       X, ldk, ldp, P, kpad, o);
  }
}

// wpb: waves (16-point tiles) per workgroup, 4 or 2; mt: 16-orbital tiles
// per wave (0: gemm_orbital_tiles' rule).  Every output element accumulates
// the same k-chunks in the same order whatever the choice, so it changes
// speed only, never a result bit.
template <typename TI, typename TO>
hipError_t launch_orbital_gemm(const TI *Ct, int ldc, int n_orb, const TI *X, int ldk, long ldp,
                               long P, int kpad, const GemmOut &o, hipStream_t st, int wpb,
                               int mt) {
  if (P <= 0 || n_orb <= 0) return hipSuccess;
  if (mt == 0) mt = gemm_orbital_tiles(n_orb, P);
  if (!gemm_orbital_tiles_valid(mt, n_orb, ldc)) return hipErrorInvalidValue;
  if (wpb == 2)
    gemm_grid<TI, TO, 2>(Ct, ldc, n_orb, X, ldk, ldp, P, kpad, o, st, mt);
  else
    gemm_grid<TI, TO, 4>(Ct, ldc, n_orb, X, ldk, ldp, P, kpad, o, st, mt);
  return hipGetLastError();
}

template hipError_t launch_orbital_gemm<double, double>(const double *, int, int, const double *,
                                                        int, long, long, int, const GemmOut &,
                                                        hipStream_t, int, int);
template hipError_t launch_orbital_gemm<float, float>(const float *, int, int, const float *, int,
                                                      long, long, int, const GemmOut &,
                                                      hipStream_t, int, int);
template hipError_t launch_orbital_gemm<double, float>(const double *, int, int, const double *,
                                                       int, long, long, int, const GemmOut &,
                                                       hipStream_t, int, int);
template hipError_t launch_orbital_gemm<float, bf16s>(const float *, int, int, const float *, int,
                                                      long, long, int, const GemmOut &,
                                                      hipStream_t, int, int);
template hipError_t launch_orbital_gemm<__hip_bfloat16, float>(const __hip_bfloat16 *, int, int,
                                                               const __hip_bfloat16 *, int, long,
                                                               long, int, const GemmOut &,
                                                               hipStream_t, int, int);

}  // namespace qmcb
