// ebe_element.h -- the stored-matrix element product y_e = K_e x_e (gfx950),
// shared by the two-phase EBE operator (fem_kernels.hip) and the patch
// operator (patch.hip).
//
// Reference: the element matrices of src/problems_elasticity.jl:203-451
// (Km += w BL' D BL, :369-375), applied to x_e.
#pragma once
#include "jfem_device.h"

namespace jfem {

// The upper blocks (i <= j) of an element matrix in row-major pair order.
template <int NPE> struct PairTable { int i[Elem<NPE>::NPAIR], j[Elem<NPE>::NPAIR]; };
template <int NPE> constexpr PairTable<NPE> make_pairs() {
    PairTable<NPE> t{};
    int p = 0;
    for (int i = 0; i < NPE; ++i)
        for (int j = i; j < NPE; ++j) { t.i[p] = i; t.j[p] = j; ++p; }
    return t;
}
template <int NPE> __host__ __device__ constexpr int pair_index(int i, int j) {
    return i * NPE - (i * (i - 1)) / 2 + (j - i);
}

// The element matrices are read once per product: non-temporal loads keep
// them from displacing x and the connectivity in the caches (at 1.05M DOFs
// the FP64 element phase went from 0.214 to 0.185 ms, FP32 from 0.105 to
// 0.090 ms).
#ifndef JFEM_EBE_NT
#define JFEM_EBE_NT 1
#endif

// Element k of the stored order: local node i at conn[i * n_el + k]; the
// upper blocks in the operator's precision as chunks of V = 16 / sizeof(T)
// values (two doubles, four floats) interleaved by element, entry q = 9 p + m
// (block p, entry m) at Ke[((q / V) n_el + k) V + q % V], so that a lane
// loads 16 bytes per instruction and the lanes of a wave read consecutive
// addresses.  f receives the nodes' terms (an output array rather than a
// callback: with a callback the compiler spilled the fully unrolled product
// to scratch).
template <int NPE, typename T>
__device__ __forceinline__ void ebe_element(int n_el, int k, const int32_t *__restrict__ conn, const T *__restrict__ Ke,
                                            const T *__restrict__ x, T (&f)[NPE][3]) {
    using EL = Elem<NPE>;
    constexpr PairTable<NPE> PT = make_pairs<NPE>();
    // K_e annihilates translations: subtracting node 0's value keeps FP32 accurate
    T u[NPE][3], t[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) t[c] = x[3 * (size_t)conn[k] + c];
#pragma unroll
    for (int i = 0; i < NPE; ++i) {
        const int nd = conn[(size_t)i * n_el + k];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            u[i][c] = x[3 * (size_t)nd + c] - t[c];
            f[i][c] = T(0);
        }
    }
    constexpr int V = 16 / sizeof(T), NQ = 9 * EL::NPAIR;
    struct alignas(16) Chunk { T v[V]; };
    const Chunk *Kc = reinterpret_cast<const Chunk *>(Ke);
    T B[9];
#pragma clang loop unroll(full)
    for (int ch = 0; ch < (NQ + V - 1) / V; ++ch) {
#if JFEM_EBE_NT
        Chunk kv;
#pragma unroll
        for (int v = 0; v < V; ++v) kv.v[v] = __builtin_nontemporal_load(&Kc[(size_t)ch * n_el + k].v[v]);
#else
        const Chunk kv = Kc[(size_t)ch * n_el + k];
#endif
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int q = ch * V + v;
            if (q < NQ) {
                B[q % 9] = kv.v[v];
                if (q % 9 == 8) {  // block p complete
                    const int p = q / 9, i = PT.i[p], j = PT.j[p];
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        f[i][c] += B[c * 3] * u[j][0] + B[c * 3 + 1] * u[j][1] + B[c * 3 + 2] * u[j][2];
                    if (i != j) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) f[j][c] += B[c] * u[i][0] + B[3 + c] * u[i][1] + B[6 + c] * u[i][2];
                    }
                }
            }
        }
    }
}

}  // namespace jfem
