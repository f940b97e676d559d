// sort.hip -- stable LSD radix sort of (key, index) pairs and an inclusive max-scan (gfx950).
//
// The localization path sorts rows by frame (frame-sorted compaction), by (frame, cell) keys
// (neighbour counts of the statistics, link search of consecutive-frame merging).  Each pass
// sorts by one 8-bit digit:
//   radix_hist_kernel     one work-group per tile of 2048 keys: LDS digit histogram, written
//                         digit-major (counts[d * ntiles + tile]);
//   scan                  exclusive scan of the 256 x ntiles counts (digit-major), so a tile's
//                         slice of digit d starts after every earlier digit and every earlier
//                         tile's keys of digit d;
//   radix_scatter_kernel  the tile again, 256 keys per round in index order: each wave finds the
//                         lanes holding its digit by 8 ballots (one per digit bit), ranks itself
//                         among the lower lanes, adds the counts of the lower waves (LDS) and the
//                         tile's running base of the digit.
// Equal digits keep their input order (stable), so the passes from the lowest digit up sort the
// keys stably; there are no atomics on the data path, so the result is deterministic.
#include <limits.h>
#include "qcs_internal.h"

namespace qcs {

namespace {

constexpr int SB = 256;                  // threads per work-group (4 waves)
constexpr int SI = 8;                    // keys per thread per tile
constexpr int STILE = SB * SI;           // keys per tile
constexpr int RADIX = 256;

template <typename K>
__global__ void __launch_bounds__(SB) radix_hist_kernel(const K *__restrict__ keys, int n, int shift, int ntiles,
                                                        int *__restrict__ counts)
{
    __shared__ int h[RADIX];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int t0 = blockIdx.x * STILE;
#pragma unroll
    for (int k = 0; k < SI; k++) {
        const int i = t0 + k * SB + tid;
        if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & (K)(RADIX - 1))], 1);
    }
    __syncthreads();
    counts[(size_t)tid * ntiles + blockIdx.x] = h[tid];
}

template <typename K>
__global__ void __launch_bounds__(SB) radix_scatter_kernel(const K *__restrict__ kin, const int *__restrict__ vin,
                                                           K *__restrict__ kout, int *__restrict__ vout, int n, int shift,
                                                           int ntiles, const int *__restrict__ offs)
{
    __shared__ int base[RADIX];
    __shared__ int wcnt[SB / 64][RADIX];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    base[tid] = offs[(size_t)tid * ntiles + blockIdx.x];
    const int t0 = blockIdx.x * STILE;
    for (int k = 0; k < SI; k++) {
#pragma unroll
        for (int w = 0; w < SB / 64; w++) wcnt[w][tid] = 0;
        __syncthreads();
        const int i = t0 + k * SB + tid;
        const bool valid = i < n;
        const K key = valid ? kin[i] : (K)0;
        const int val = valid ? vin[i] : 0;
        const int d = (int)((key >> shift) & (K)(RADIX - 1));
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const unsigned long long bb = __ballot((d >> b) & 1);
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        if (valid && rank == 0) wcnt[wv][d] = __popcll(m);   // one lane per digit present in the wave
        __syncthreads();
        if (valid) {
            int pos = base[d] + rank;
            for (int w = 0; w < wv; w++) pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < SB / 64; w++) add += wcnt[w][tid];
        base[tid] += add;
    }
}

// inclusive max-scan: per-tile scan, scan of the tile maxima, then the carry-in
constexpr int MB = 256, MI = 16, MTILE = MB * MI;

__device__ __forceinline__ int block_inclusive_max(int v, int *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < MB; off <<= 1) {
        const int t = (tid >= off) ? sh[tid - off] : INT_MIN;
        __syncthreads();
        sh[tid] = sh[tid] > t ? sh[tid] : t;
        __syncthreads();
    }
    const int r = sh[tid];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(MB) max_scan_tiles_kernel(const int *__restrict__ in, int n, int *__restrict__ out,
                                                           int *__restrict__ tile_max)
{
    __shared__ int sh[MB];
    const int base = blockIdx.x * MTILE + threadIdx.x * MI;
    int v[MI], run = INT_MIN;
#pragma unroll
    for (int k = 0; k < MI; k++) {
        v[k] = (base + k < n) ? in[base + k] : INT_MIN;
        run = run > v[k] ? run : v[k];
    }
    const int incl = block_inclusive_max(run, sh);   // sh keeps every thread's inclusive maximum
    int carry = (threadIdx.x > 0) ? sh[threadIdx.x - 1] : INT_MIN;
#pragma unroll
    for (int k = 0; k < MI; k++) {
        carry = carry > v[k] ? carry : v[k];
        if (base + k < n) out[base + k] = carry;
    }
    if (threadIdx.x == MB - 1) tile_max[blockIdx.x] = incl;
}

__global__ void __launch_bounds__(MB) max_scan_sums_kernel(int *__restrict__ tile_max, int ntiles)
{
    __shared__ int sh[MB];
    const int base = threadIdx.x * MI;
    int v[MI], run = INT_MIN;
#pragma unroll
    for (int k = 0; k < MI; k++) {
        v[k] = (base + k < ntiles) ? tile_max[base + k] : INT_MIN;
        run = run > v[k] ? run : v[k];
    }
    (void)block_inclusive_max(run, sh);   // sh keeps every thread's inclusive maximum
    int carry = (threadIdx.x > 0) ? sh[threadIdx.x - 1] : INT_MIN;
    // exclusive carry per tile: the maximum of the tiles before it
#pragma unroll
    for (int k = 0; k < MI; k++) {
        const int x = v[k];
        if (base + k < ntiles) tile_max[base + k] = carry;
        carry = carry > x ? carry : x;
    }
}

__global__ void __launch_bounds__(256) max_scan_add_kernel(int *__restrict__ out, int n, const int *__restrict__ tile_max)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = tile_max[i / MTILE];
    out[i] = out[i] > c ? out[i] : c;
}

template <typename K>
int radix_sort_impl(K *k0, int *v0, K *k1, int *v1, int n, int bits, int *tmp, hipStream_t s, K **kres, int **vres)
{
    *kres = k0;
    *vres = v0;
    if (n <= 0 || bits <= 0) return QCS_OK;
    const int ntiles = (n + STILE - 1) / STILE;
    const long long ncount = (long long)RADIX * ntiles;
    if (ncount > 16LL * 1024 * 1024) return fail(QCS_ERR_ARG, "sort of %d keys exceeds the supported size", n);
    int *counts = tmp, *offs = tmp + ncount, *tsum = offs + ncount;
    K *ki = k0, *ko = k1;
    int *vi = v0, *vo = v1;
    const int maxbits = (int)(8 * sizeof(K));
    for (int shift = 0; shift < bits && shift < maxbits; shift += 8) {
        hipLaunchKernelGGL((radix_hist_kernel<K>), dim3(ntiles), dim3(SB), 0, s, ki, n, shift, ntiles, counts);
        QCS_HIP(hipGetLastError());
        int rc = scan_exclusive(counts, (int)ncount, offs, tsum, s);
        if (rc) return rc;
        hipLaunchKernelGGL((radix_scatter_kernel<K>), dim3(ntiles), dim3(SB), 0, s, ki, vi, ko, vo, n, shift, ntiles, offs);
        QCS_HIP(hipGetLastError());
        K *tk = ki; ki = ko; ko = tk;
        int *tv = vi; vi = vo; vo = tv;
    }
    *kres = ki;
    *vres = vi;
    return QCS_OK;
}

}  // namespace

size_t radix_sort_tmp_ints(int n)
{
    const size_t ntiles = ((size_t)(n > 0 ? n : 1) + STILE - 1) / STILE;
    const size_t ncount = (size_t)RADIX * ntiles;
    return 2 * ncount + (size_t)scan_tile_count(ncount) + 1;
}

int radix_sort_pairs_u32(unsigned *k0, int *v0, unsigned *k1, int *v1, int n, int bits, int *tmp, hipStream_t s,
                         unsigned **kres, int **vres)
{
    return radix_sort_impl<unsigned>(k0, v0, k1, v1, n, bits, tmp, s, kres, vres);
}

int radix_sort_pairs_u64(unsigned long long *k0, int *v0, unsigned long long *k1, int *v1, int n, int bits, int *tmp,
                         hipStream_t s, unsigned long long **kres, int **vres)
{
    return radix_sort_impl<unsigned long long>(k0, v0, k1, v1, n, bits, tmp, s, kres, vres);
}

size_t max_scan_tmp_ints(int n) { return (size_t)(((size_t)(n > 0 ? n : 1) + MTILE - 1) / MTILE) + 1; }

int scan_inclusive_max(const int *in, int n, int *out, int *tmp, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const int ntiles = (n + MTILE - 1) / MTILE;
    if (ntiles > MTILE) return fail(QCS_ERR_ARG, "max-scan of %d items exceeds the supported size", n);
    hipLaunchKernelGGL(max_scan_tiles_kernel, dim3(ntiles), dim3(MB), 0, s, in, n, out, tmp);
    hipLaunchKernelGGL(max_scan_sums_kernel, dim3(1), dim3(MB), 0, s, tmp, ntiles);
    hipLaunchKernelGGL(max_scan_add_kernel, dim3((n + 255) / 256), dim3(256), 0, s, out, n, tmp);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

}  // namespace qcs
