// jfem_device.h -- device helpers shared by the kernel files.
//
// Reductions are deterministic and free of atomics: a producer kernel writes
// one partial per block (fixed xor-butterfly inside each wave, waves summed in
// index order), and every block of the consuming kernel sums the same
// partials in the same fixed order, so all blocks obtain bit-identical
// scalars without a grid-wide synchronisation or a memory fence.
#pragma once

#include "jfem_internal.h"

namespace jfem {

constexpr int kBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block, returned to every thread.
template <int NV> __device__ __forceinline__ void block_sum_all(double (&v)[NV]) {
    __shared__ double sh[NV][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k][wid] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += sh[k][w];
        v[k] = s;
    }
    __syncthreads();
}

// Block sums of v[] -> part[k*kMaxPart + blockIdx.x].  Called by all threads.
template <int NV> __device__ __forceinline__ void write_partials(double (&v)[NV], double *part, int slot) {
    block_sum_all<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) part[k * kMaxPart + slot] = v[k];
    }
}
template <int NV> __device__ __forceinline__ void write_partials(double (&v)[NV], double *part) {
    write_partials<NV>(v, part, (int)blockIdx.x);
}

// Sum of npart partials per value, fixed order, returned to every thread.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *__restrict__ part, int npart, double (&out)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < npart; i += blockDim.x) s += part[k * kMaxPart + i];
        out[k] = s;
    }
    block_sum_all<NV>(out);
}

// The same sums formed by one wave alone, with no barrier, so that a kernel
// can take them while its other loads are in flight: lane l loads partials
// l, l + 64, ... (all issued before the first add), adds them in that order,
// then the xor butterfly.  Every wave of every block obtains the same bits
// (the butterfly's additions are commutative pairs), which differ from
// sum_partials' only in the order of the additions.
template <int NV> struct WavePartials {
    static constexpr int kPer = kMaxPart / 64;
    double v[NV][kPer];
    __device__ __forceinline__ void load(const double *__restrict__ part, int npart) {
        const int lane = (int)threadIdx.x & 63;
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int j = 0; j < kPer; ++j) {
                const int i = lane + 64 * j;
                v[k][j] = i < npart ? part[k * kMaxPart + i] : 0.0;
            }
    }
    __device__ __forceinline__ void sum(double (&out)[NV]) const {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < kPer; ++j) s += v[k][j];
            out[k] = wave_sum(s);
        }
    }
};

__device__ __forceinline__ bool finite(double x) { return x - x == 0.0; }

// Loads at a 32-bit byte offset from a uniform base pointer, so that the
// compiler addresses them as SGPR base + VGPR offset with no 64-bit address
// arithmetic per lane (every array indexed this way is below 4 GiB).
template <typename T> struct Trip { T a, b, c; };
template <typename V, typename T>
__device__ __forceinline__ V load_at(const T *__restrict__ base, uint32_t byte_off) {
    return *reinterpret_cast<const V *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <typename V, typename T>
__device__ __forceinline__ void store_at(T *__restrict__ base, uint32_t byte_off, const V &v) {
    *reinterpret_cast<V *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

// A node's three values (and a node's 3 x 3 block) as one vector access
// each, instead of one per component.
template <typename T> struct Blk9 { T v[9]; };
template <typename T> __device__ __forceinline__ Trip<T> ld3(const T *__restrict__ p, size_t node) {
    return reinterpret_cast<const Trip<T> *>(p)[node];
}
template <typename T> __device__ __forceinline__ void st3(T *__restrict__ p, size_t node, T a, T b, T c) {
    reinterpret_cast<Trip<T> *>(p)[node] = Trip<T>{a, b, c};
}
template <typename T> __device__ __forceinline__ Blk9<T> ld9(const T *__restrict__ p, size_t node) {
    return reinterpret_cast<const Blk9<T> *>(p)[node];
}

// Wave-local barrier: the LDS a wave reads after it was written by other
// lanes of the same wave only.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// dN/dxi at the Gauss points as constant tables for runtime (q, i) indexing
// (compile-time indexing uses Elem<NPE>::dN directly).
template <int NPE> struct DNTable { double v[Elem<NPE>::NGP][NPE][3]; };
template <int NPE> constexpr DNTable<NPE> make_dn() {
    DNTable<NPE> t{};
    for (int q = 0; q < Elem<NPE>::NGP; ++q)
        for (int i = 0; i < NPE; ++i)
            for (int c = 0; c < 3; ++c) t.v[q][i][c] = Elem<NPE>::dN(q, i, c);
    return t;
}
static __constant__ DNTable<10> c_dn10 = make_dn<10>();
static __constant__ DNTable<4> c_dn4 = make_dn<4>();
static __constant__ DNTable<8> c_dn8 = make_dn<8>();

template <int NPE> __device__ __forceinline__ double dn_rt(int q, int i, int c);
template <> __device__ __forceinline__ double dn_rt<10>(int q, int i, int c) { return c_dn10.v[q][i][c]; }
template <> __device__ __forceinline__ double dn_rt<4>(int q, int i, int c) { return c_dn4.v[q][i][c]; }
template <> __device__ __forceinline__ double dn_rt<8>(int q, int i, int c) { return c_dn8.v[q][i][c]; }

// Grid size for a grid-stride producer kernel: enough blocks for `work`
// items at `per_block` items per block, capped at kMaxPart.
// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD
// b % 8), and each XCD has its own L2.  xcd_block() renumbers the blocks of a
// grid whose size is a multiple of 8 (xcd_grid) so that every XCD works on one
// contiguous eighth of the logical blocks: neighbouring elements or rows, and
// the parts of x they read, then share an L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int q = (int)gridDim.x / kXcds;
    return ((int)blockIdx.x % kXcds) * q + (int)blockIdx.x / kXcds;
}
__host__ __device__ inline unsigned xcd_grid(unsigned nb) { return (nb + kXcds - 1) / kXcds * kXcds; }

// Surface elements: Tri6 / Tri3 / Quad4 basis (src/basis/lagrange_generated.jl:
// 89-171) and their rules (Tri 3-point and 1-point, Quad 2x2 Gauss).
__device__ inline void surf_basis(int nfn, double u, double v, double *N, double (*dN)[2]) {
    if (nfn == 6) {
        N[0] = 1 - 3.0 * u - 3.0 * v + 2.0 * u * u + 4.0 * u * v + 2.0 * v * v;
        N[1] = -u + 2.0 * u * u;
        N[2] = -v + 2.0 * v * v;
        N[3] = 4.0 * u - 4.0 * u * u - 4.0 * u * v;
        N[4] = 4.0 * u * v;
        N[5] = 4.0 * v - 4.0 * u * v - 4.0 * v * v;
        dN[0][0] = -3.0 + 4.0 * u + 4.0 * v; dN[0][1] = -3.0 + 4.0 * u + 4.0 * v;
        dN[1][0] = -1.0 + 4.0 * u; dN[1][1] = 0.0;
        dN[2][0] = 0.0; dN[2][1] = -1.0 + 4.0 * v;
        dN[3][0] = 4.0 - 8.0 * u - 4.0 * v; dN[3][1] = -4.0 * u;
        dN[4][0] = 4.0 * v; dN[4][1] = 4.0 * u;
        dN[5][0] = -4.0 * v; dN[5][1] = 4.0 - 4.0 * u - 8.0 * v;
    } else if (nfn == 3) {
        N[0] = 1 - u - v; N[1] = u; N[2] = v;
        dN[0][0] = -1; dN[0][1] = -1; dN[1][0] = 1; dN[1][1] = 0; dN[2][0] = 0; dN[2][1] = 1;
    } else {
        const double r[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        for (int i = 0; i < 4; ++i) {
            const double a = 1 + r[i][0] * u, b = 1 + r[i][1] * v;
            N[i] = 0.25 * a * b;
            dN[i][0] = 0.25 * r[i][0] * b;
            dN[i][1] = 0.25 * a * r[i][1];
        }
    }
}

__device__ inline int surf_rule(int nfn, double (&w)[4], double (&pu)[4], double (&pv)[4]) {
    if (nfn == 6) {
        w[0] = w[1] = w[2] = 1.0 / 6;
        pu[0] = 1.0 / 6; pv[0] = 1.0 / 6;
        pu[1] = 2.0 / 3; pv[1] = 1.0 / 6;
        pu[2] = 1.0 / 6; pv[2] = 2.0 / 3;
        return 3;
    }
    if (nfn == 3) {
        w[0] = 0.5; pu[0] = pv[0] = 1.0 / 3;
        return 1;
    }
    const double a = 0.57735026918962576451;  // 1/sqrt(3)
    w[0] = w[1] = w[2] = w[3] = 1.0;
    pu[0] = -a; pv[0] = -a; pu[1] = a; pv[1] = -a; pu[2] = -a; pv[2] = a; pu[3] = a; pv[3] = a;
    return 4;
}

inline unsigned capped_grid(int64_t work, int per_block) {
    int64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > kMaxPart) g = kMaxPart;
    return (unsigned)g;
}

}  // namespace jfem
