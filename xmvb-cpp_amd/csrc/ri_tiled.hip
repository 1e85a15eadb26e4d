// RI one-index transform over aux-tiled factors: every packed element of B is read from HBM once.
//
// transform_ri_one_index (integrals/ao/ri_transform.cpp:10-37) forms, per auxiliary index A,
//   T[A, q, mu] = sum_nu B_A[mu, nu] Cn[nu, q]
// with B_A symmetric and stored packed (pair (mu, nu) at hi(hi+1)/2 + lo). A packed element
// (mu > nu) feeds two rows of T: T[., mu] through Cn[nu, .] and T[., nu] through Cn[mu, .].
//
// Layout (written at upload): the factors of kRiTileAux consecutive aux indices form one tile,
// tiles one after another, inside a tile pair-major with the tile's aux indices contiguous:
//   Bt[(tile * npair + pair) * kRiTileAux + a]          (aux A = tile * kRiTileAux + a)
// so the packed rows of a tile are contiguous runs of HBM (2 KB per 32 pairs in FP64).
//
// One workgroup (256 threads) owns one aux tile of one slot and ALL of its AO rows: thread
// (a = t % 8, r0 = t / 8) accumulates T[A, :, r0 + 32 k] for every 32-row band k in registers. The
// packed triangle is walked in 32 x 32 blocks (k, j), j <= k; block (k, j) is staged in LDS once
// and serves both of its uses: the band-k rows (row part, sum over nu in band j) and the band-j
// rows (column part, sum over mu in band k). Every thread has exactly one row in each band, so
// every thread does the same work on every block (32 elements for a diagonal block, 64 for an
// off-diagonal one) and no accumulation crosses threads.
//
// Storage variants: FP64, and compensated FP32 (hi = fp32(B) plus an 8-bit correction k of its
// rounding error in units of ulp(hi) / 256: B = hi + k 2^(e(hi) - 158), exact in FP64, 2^-32
// relative), decoded to FP64 when a block is staged; the arithmetic is FP64 throughout.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

#include "device.hpp"
#include "kernels.hpp"

namespace xmvb_amd {

namespace {

constexpr int kBand = 32;                       // AO rows per band (= threads / kRiTileAux)
constexpr int kThreads = kBand * kRiTileAux;    // 256
constexpr int kRowStride = kBand * kRiTileAux + 8;  // LDS doubles per staged row (64 B of padding)
constexpr int kTileDoubles = kBand * kRowStride;

// decodes one compensated element: hi + k * 2^(e - 158), e the biased exponent of hi (at least 1)
__device__ __forceinline__ double decode_fp32c(float hi, int k) {
    unsigned e = (__float_as_uint(hi) >> 23) & 0xffu;
    e = e ? e : 1u;
    const double scale = __hiloint2double((int)((e + 1023u - 158u) << 20), 0);
    return fma((double)k, scale, (double)hi);
}

// Staging of block (k, j) of one slot's tile: thread t carries kChunks groups of 4 consecutive
// doubles of the block (row mu' = group / 64, first element (group % 64) * 4 of the row's 256),
// loaded into registers while the previous block is being used, then decoded into the LDS tile.
template <bool Comp>
struct Stage {
    static constexpr int kGroups = kBand * kBand * kRiTileAux / 4 / kThreads;  // 8
    double v[kGroups][4];

    // row mu' of block (k, j) holds pairs (32k + mu', 32j + nu'), nu' < 32 (nu' <= mu' on the diagonal)
    __device__ __forceinline__ void load(const SlotDesc& sd, int tile, int k, int j) {
        const int n_bf = sd.n_bf;
        const long long npair = (long long)n_bf * (n_bf + 1) / 2;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const int grp = (int)threadIdx.x + kThreads * g;
            const int mu_l = grp >> 6, off = (grp & 63) * 4;  // element offset in the row (nu' * 8 + a)
            const int mu = kBand * k + mu_l, nu = kBand * j + (off >> 3);
            const bool ok = mu < n_bf && nu <= mu;  // nu <= mu: inside the packed row (diagonal block)
            const long long pair = (long long)mu * (mu + 1) / 2 + (ok ? nu : 0);
            const long long at = ((long long)tile * npair + (ok ? pair : 0)) * kRiTileAux + (off & 7);
            if constexpr (Comp) {
                float4 h = make_float4(0.f, 0.f, 0.f, 0.f);
                int lo = 0;
                if (ok) {
                    const f32x4 hv = *reinterpret_cast<gptr<const f32x4>>(to_global(sd.B32 + at));
                    h = make_float4(hv[0], hv[1], hv[2], hv[3]);
                    lo = *reinterpret_cast<gptr<const int>>(to_global(sd.Blo + at));
                }
                v[g][0] = decode_fp32c(h.x, (int)(signed char)(lo & 0xff));
                v[g][1] = decode_fp32c(h.y, (int)(signed char)((lo >> 8) & 0xff));
                v[g][2] = decode_fp32c(h.z, (int)(signed char)((lo >> 16) & 0xff));
                v[g][3] = decode_fp32c(h.w, (int)(signed char)((lo >> 24) & 0xff));
            } else {
                double2 p0 = make_double2(0.0, 0.0), p1 = make_double2(0.0, 0.0);
                if (ok) {
                    const gptr<const f64x2> src = reinterpret_cast<gptr<const f64x2>>(to_global(sd.B + at));
                    const f64x2 a0 = src[0], a1 = src[1];
                    p0 = make_double2(a0[0], a0[1]);
                    p1 = make_double2(a1[0], a1[1]);
                }
                v[g][0] = p0.x;
                v[g][1] = p0.y;
                v[g][2] = p1.x;
                v[g][3] = p1.y;
            }
        }
    }

    __device__ __forceinline__ void store(double* tile_lds) const {
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const int grp = (int)threadIdx.x + kThreads * g;
            double2* dst = reinterpret_cast<double2*>(tile_lds + (grp >> 6) * kRowStride + (grp & 63) * 4);
            dst[0] = make_double2(v[g][0], v[g][1]);
            dst[1] = make_double2(v[g][2], v[g][3]);
        }
    }
};

// Cn as [nu][kCnStride] (SlotDesc::CnT, written by k_orbital_prep, zeros past n and n_bf), read
// through the constant address space: every lane of a wavefront reads the same element, so the
// loads go through the scalar unit and the values reach the FP64 FMAs as scalar operands, leaving
// the LDS to the staged block alone.
using CPtr = const __attribute__((address_space(4))) double*;

// acc_row[q] += sum_nu' L[r0][nu'] c_j[nu'][q] (band-k row over band-j columns) and, off the
// diagonal, acc_col[q] += sum_mu' L[mu'][r0] c_k[mu'][q] (band-j row over band-k rows). On the
// diagonal block the symmetric element (r0, nu') is L[r0][nu'] for nu' <= r0, else L[nu'][r0].
// acc_row[q] += sum_nu' L[r0][nu'] c_j[nu'][q] (band-k row over band-j columns) and, off the
// diagonal, acc_col[q] += sum_mu' L[mu'][r0] c_k[mu'][q] (band-j row over band-k rows). On the
// diagonal block the symmetric element (r0, nu') is L[r0][nu'] for nu' <= r0, else L[nu'][r0].
template <int NQ, bool Diag>
__device__ __forceinline__ void use_block(const double* __restrict__ L, CPtr cj, CPtr ck, int r0, int a,
                                          double (&acc_row)[NQ], double (&acc_col)[NQ]) {
    const double* row = L + r0 * kRowStride + a;  // L[r0][nu'] at row[nu' * 8]
    const double* col = L + r0 * kRiTileAux + a;  // L[mu'][r0] at col[mu' * kRowStride]
    if constexpr (Diag) {
#pragma unroll 4
        for (int t = 0; t < kBand; ++t) {
            const double b = L[(t <= r0 ? r0 * kRowStride + t * kRiTileAux : t * kRowStride + r0 * kRiTileAux) + a];
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc_row[q] = fma(b, cj[t * kCnStride + q], acc_row[q]);
        }
    } else {
#pragma unroll 4
        for (int t = 0; t < kBand; ++t) {
            const double b = row[t * kRiTileAux];
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc_row[q] = fma(b, cj[t * kCnStride + q], acc_row[q]);
        }
#pragma unroll 4
        for (int t = 0; t < kBand; ++t) {
            const double b = col[t * kRowStride];
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc_col[q] = fma(b, ck[t * kCnStride + q], acc_col[q]);
        }
    }
}

// grid (tiles, 1, slots); NB: bands covering the launch's largest n_bf; NQ: n rounded up to even.
template <int NQ, int NB, bool Comp>
__global__ void __launch_bounds__(kThreads) k_ri_one_index_tiled(const SlotDesc* __restrict__ slots, int n) {
    const SlotDesc& sd = slots[blockIdx.z];
    const int n_bf = sd.n_bf, n_aux = sd.n_aux;
    const int tile = blockIdx.x;
    if (!sd.b_tiled || tile * kRiTileAux >= n_aux) return;  // block-uniform
    extern __shared__ double lds[];
    double* L = lds;  // staged block, kTileDoubles
    const CPtr cl = (CPtr)sd.CnT;
    const int a = threadIdx.x & (kRiTileAux - 1), r0 = threadIdx.x / kRiTileAux;
    const int nb = (n_bf + kBand - 1) / kBand;  // bands of this slot (<= NB)
    double acc[NB][NQ];
#pragma unroll
    for (int k = 0; k < NB; ++k)
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[k][q] = 0.0;
    Stage<Comp> st;
    st.load(sd, tile, 0, 0);
    // blocks (k, j), j <= k, in order; compile-time k and j keep every accumulator in registers
    auto block = [&]<int K, int J>() {
        if (K >= nb) return;  // block-uniform
        __syncthreads();  // the previous block is no longer read
        st.store(L);
        __syncthreads();
        constexpr int K1 = J < K ? K : K + 1, J1 = J < K ? J + 1 : 0;
        if (K1 < nb) st.load(sd, tile, K1, J1);  // the next block, loaded while this one is used
        if constexpr (J == K) {
            double unused[NQ];
            use_block<NQ, true>(L, cl + kBand * K * kCnStride, cl, r0, a, acc[K], unused);
        } else {
            use_block<NQ, false>(L, cl + kBand * J * kCnStride, cl + kBand * K * kCnStride, r0, a, acc[K], acc[J]);
        }
    };
    auto band = [&]<int K>() {
        [&]<int... Js>(std::integer_sequence<int, Js...>) {
            (block.template operator()<K, Js>(), ...);
        }(std::make_integer_sequence<int, K + 1>{});
    };
    [&]<int... Ks>(std::integer_sequence<int, Ks...>) {
        (band.template operator()<Ks>(), ...);
    }(std::make_integer_sequence<int, NB>{});
    const int A = tile * kRiTileAux + a;
    if (A >= n_aux) return;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int mu = kBand * k + r0;
        if (mu >= n_bf) break;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            if (q < n) to_global(sd.T)[(size_t)A + (size_t)n_aux * (q + (size_t)n * mu)] = acc[k][q];
    }
}

template <int NQ, int NB, bool Comp>
hipError_t launch_tiled(bool run, hipStream_t st, dim3 grid, const SlotDesc* slots, int n) {
    const size_t lds = (size_t)kTileDoubles * sizeof(double);
    if (!run)
        return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_ri_one_index_tiled<NQ, NB, Comp>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_ri_one_index_tiled<NQ, NB, Comp>), grid, dim3(kThreads), lds, st, slots, n);
    return hipSuccess;
}

template <int NQ, bool Comp>
hipError_t dispatch_nb(bool run, int nb, hipStream_t st, dim3 grid, const SlotDesc* slots, int n) {
    switch (nb) {
        case 1: return launch_tiled<NQ, 1, Comp>(run, st, grid, slots, n);
        case 2: return launch_tiled<NQ, 2, Comp>(run, st, grid, slots, n);
        case 3: return launch_tiled<NQ, 3, Comp>(run, st, grid, slots, n);
        case 4: return launch_tiled<NQ, 4, Comp>(run, st, grid, slots, n);
        case 5: return launch_tiled<NQ, 5, Comp>(run, st, grid, slots, n);
        case 6: return launch_tiled<NQ, 6, Comp>(run, st, grid, slots, n);
        default: return launch_tiled<NQ, 7, Comp>(run, st, grid, slots, n);
    }
}

template <bool Comp>
hipError_t dispatch(bool run, int n, int nb, hipStream_t st, dim3 grid, const SlotDesc* slots) {
    switch ((n + 1) / 2) {
        case 1: return dispatch_nb<2, Comp>(run, nb, st, grid, slots, n);
        case 2: return dispatch_nb<4, Comp>(run, nb, st, grid, slots, n);
        case 3: return dispatch_nb<6, Comp>(run, nb, st, grid, slots, n);
        case 4: return dispatch_nb<8, Comp>(run, nb, st, grid, slots, n);
        case 5: return dispatch_nb<10, Comp>(run, nb, st, grid, slots, n);
        case 6: return dispatch_nb<12, Comp>(run, nb, st, grid, slots, n);
        default: return dispatch_nb<14, Comp>(run, nb, st, grid, slots, n);
    }
}

// compensated encoding of one FP64 value: hi = fp32(b), k = round((b - hi) / 2^(e(hi) - 158))
__device__ __forceinline__ void encode_fp32c(double b, float& hi, signed char& k) {
    hi = __double2float_rn(b);
    unsigned e = (__float_as_uint(hi) >> 23) & 0xffu;
    e = e ? e : 1u;
    const double inv_scale = __hiloint2double((int)((158u - e + 1023u) << 20), 0);  // 2^(158 - e)
    const double r = rint((b - (double)hi) * inv_scale);
    k = (signed char)fmin(127.0, fmax(-128.0, r));
}

// one thread per element of the destination layout
__global__ void __launch_bounds__(256) k_ri_relayout(const double* __restrict__ flat, int n_aux, long long npair,
                                                     int tiled, long long total, double* __restrict__ B,
                                                     float* __restrict__ B32, signed char* __restrict__ Blo) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int A;
    long long pair;
    if (tiled) {
        const long long tp = i / kRiTileAux;  // tile * npair + pair
        A = (int)(tp / npair) * kRiTileAux + (int)(i % kRiTileAux);
        pair = tp % npair;
    } else {
        A = (int)(i % n_aux);
        pair = i / n_aux;
    }
    const double b = A < n_aux ? flat[(size_t)A + (size_t)n_aux * pair] : 0.0;
    if (B) {
        B[i] = b;
    } else {
        float hi;
        signed char k;
        encode_fp32c(b, hi, k);
        B32[i] = hi;
        Blo[i] = k;
    }
}

}  // namespace

hipError_t configure_ri_tiled() {
    for (int n = 2; n <= MAX_ACT; n += 2)
        for (int nb = 1; nb <= kRiTiledMaxBands; ++nb)
            for (const bool comp : {false, true}) {
                const hipError_t e = comp ? dispatch<true>(false, n, nb, nullptr, dim3(), nullptr)
                                          : dispatch<false>(false, n, nb, nullptr, dim3(), nullptr);
                if (e != hipSuccess) return e;
            }
    return hipSuccess;
}

void launch_ri_relayout(hipStream_t st, const double* flat, int n_aux, long long npair, bool tiled, double* B,
                        float* B32, signed char* Blo) {
    const long long aux = tiled ? (long long)(n_aux + kRiTileAux - 1) / kRiTileAux * kRiTileAux : n_aux;
    const long long total = aux * npair;
    hipLaunchKernelGGL(k_ri_relayout, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, flat, n_aux, npair,
                       tiled ? 1 : 0, total, B, B32, Blo);
}

void launch_ri_one_index_tiled(hipStream_t st, const SlotDesc* slots, int count, int n, int max_n_bf, int max_n_aux,
                               bool compensated) {
    const int nb = (max_n_bf + kBand - 1) / kBand;
    const dim3 grid((unsigned)((max_n_aux + kRiTileAux - 1) / kRiTileAux), 1, (unsigned)count);
    if (compensated) (void)dispatch<true>(true, n, nb, st, grid, slots);
    else (void)dispatch<false>(true, n, nb, st, grid, slots);
}

}  // namespace xmvb_amd
