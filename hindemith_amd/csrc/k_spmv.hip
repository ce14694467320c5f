// a3 SpmvCSR for gfx950.
//
// The reference runs one work-item per row (SpmvCSR.cpp:32-38), which leaves
// the data/indices reads uncoalesced and serialises long rows.  Here a row is
// owned by a sub-wave of L lanes (L = 1..64 dividing the 64-wide wavefront):
// the lanes stride the row's entries together, so a sub-wave reads L
// consecutive (value, index) pairs per step, and the partial sums are folded
// with cross-lane shuffles inside the sub-wave.  For irregular row lengths the
// rows are binned by length and each bin gets its own L.
#include "hmx_common.h"

#include <algorithm>
#include <vector>

#pragma clang fp contract(off)

namespace hmx {

template <int L>
__device__ inline float subwave_sum(float v) {
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, L);
    return v;
}

// rows: optional row list (binned launch); nrows rows handled.
template <int L>
__global__ __launch_bounds__(256) void csr_vector_kernel(int nrows, const int *__restrict__ rows,
                                                         const int *__restrict__ indptr,
                                                         const int *__restrict__ indices,
                                                         const float *__restrict__ data,
                                                         const float *__restrict__ x,
                                                         float *__restrict__ y) {
    const int lane = threadIdx.x & (L - 1);
    const int64_t sub = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / L;
    const int64_t nsub = ((int64_t)gridDim.x * blockDim.x) / L;
    // Whole sub-waves iterate together, so the shuffles below always see
    // their full sub-wave active.
    for (int64_t s = sub; s < nrows; s += nsub) {
        const int row = rows ? rows[s] : (int)s;
        const int beg = indptr[row], end = indptr[row + 1];
        float acc = 0.0f;
        for (int j = beg + lane; j < end; j += L) acc = fmaf(data[j], x[indices[j]], acc);
        acc = subwave_sum<L>(acc);
        if (lane == 0) y[row] = acc;
    }
}

static int launch_csr_L(int L, int nrows, const int *rows, const int *indptr,
                        const int *indices, const float *data, const float *x, float *y,
                        hipStream_t s) {
    if (nrows <= 0) return HMX_OK;
    int64_t threads = (int64_t)nrows * L;
    int64_t blocks = (threads + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    dim3 g((unsigned)blocks), b(256);
    switch (L) {
        case 1: hipLaunchKernelGGL(csr_vector_kernel<1>, g, b, 0, s, nrows, rows, indptr, indices, data, x, y); break;
        case 2: hipLaunchKernelGGL(csr_vector_kernel<2>, g, b, 0, s, nrows, rows, indptr, indices, data, x, y); break;
        case 4: hipLaunchKernelGGL(csr_vector_kernel<4>, g, b, 0, s, nrows, rows, indptr, indices, data, x, y); break;
        case 8: hipLaunchKernelGGL(csr_vector_kernel<8>, g, b, 0, s, nrows, rows, indptr, indices, data, x, y); break;
        case 16: hipLaunchKernelGGL(csr_vector_kernel<16>, g, b, 0, s, nrows, rows, indptr, indices, data, x, y); break;
        case 32: hipLaunchKernelGGL(csr_vector_kernel<32>, g, b, 0, s, nrows, rows, indptr, indices, data, x, y); break;
        case 64: hipLaunchKernelGGL(csr_vector_kernel<64>, g, b, 0, s, nrows, rows, indptr, indices, data, x, y); break;
        default: return fail(HMX_ERR_INVALID, "csr: lanes_per_row must be a power of two <= 64");
    }
    HMX_LAUNCH_CHECK("csr_vector");
    return HMX_OK;
}

// One row per lane, staged: a wavefront takes 64 consecutive rows, whose
// entries are one contiguous span of indices / data.  The wave reads the span
// coalesced (lane l takes entries l, l + 64, ...), gathers x for each entry,
// and puts the (data, x) pairs in LDS; then each lane runs its own row's
// chain acc = fma(data, x, acc) from LDS in entry order -- the reference's
// per-row sequential chain (SpmvCSR.cpp:32-38), so the result is the one
// lane's bit for bit.  The per-entry read of the direct form waits for its
// index before its x gather, entry after entry; here a lane's gathers are
// independent and in flight together.  A wave whose span exceeds its LDS
// share (CSRW_CAP entries) runs the direct loop.  Waves synchronise only with
// themselves (no work-group barrier), so a wave past the last row just exits.
constexpr int CSRW_NT = 256, CSRW_CAP = 640;

__global__ __launch_bounds__(CSRW_NT) void csr_wave_rows_kernel(int nrows,
                                                                const int *__restrict__ indptr,
                                                                const int *__restrict__ indices,
                                                                const float *__restrict__ data,
                                                                const float *__restrict__ x,
                                                                float *__restrict__ y) {
    __shared__ float2 stage[CSRW_NT / 64][CSRW_CAP];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * (CSRW_NT / 64) + wave) * 64;
    if (row0 >= nrows) return;
    const int64_t row = row0 + lane;
    const bool has = row < nrows;
    const int64_t last = row0 + 64 < nrows ? row0 + 64 : nrows;
    const int base = indptr[row0];
    const int span = indptr[last] - base;  // the same value in every lane
    const int beg = has ? indptr[row] : 0, end = has ? indptr[row + 1] : 0;
    float acc = 0.0f;
    if (span <= CSRW_CAP) {
        float2 *S = stage[wave];
#pragma unroll 5
        for (int k = lane; k < span; k += 64) {
            const int j = base + k;
            S[k] = make_float2(data[j], x[indices[j]]);
        }
        // the wave's LDS writes before its reads (a wavefront's LDS accesses
        // are performed in order; the fences keep the compiler's order)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int j = beg; j < end; ++j) {
            const float2 q = S[j - base];
            acc = fmaf(q.x, q.y, acc);
        }
    } else {
        for (int j = beg; j < end; ++j) acc = fmaf(data[j], x[indices[j]], acc);
    }
    if (has) y[row] = acc;
}

// Lanes per row from the mean row length: short rows (<= 8 nnz) go one row
// per lane, the reference's own sequential chain, staged through LDS by
// csr_wave_rows_kernel (2048^2 5-point operator: 55.6 us for the direct
// one-lane loop, against 94 us at 8 lanes); longer rows get the largest
// power of two <= mean / 2, at most 64.
int csr_auto_lanes(int M, int64_t nnz) {
    const double mean = M > 0 ? (double)nnz / M : 0.0;
    if (mean <= 8.0) return 1;
    int L = 1;
    while (L < 64 && 2 * L <= mean / 2) L *= 2;
    return L;
}

int launch_csr(int M, int64_t nnz, const int *indptr, const int *indices, const float *data,
               const float *x, float *y, int L, hipStream_t s) {
    if (L <= 0) L = csr_auto_lanes(M, nnz);
    if (L == 1 && M > 0) {
        const int64_t blocks = ((int64_t)M + CSRW_NT - 1) / CSRW_NT;
        if (blocks <= 0x7fffffff) {
            hipLaunchKernelGGL(csr_wave_rows_kernel, dim3((unsigned)blocks), dim3(CSRW_NT), 0, s,
                               M, indptr, indices, data, x, y);
            HMX_LAUNCH_CHECK("csr_wave_rows");
            return HMX_OK;
        }
    }
    return launch_csr_L(L, M, nullptr, indptr, indices, data, x, y, s);
}

// Length classes of the binned launch and the sub-wave width used for each.
static const int kBinMax[4] = {4, 16, 64, 1 << 30};
static const int kBinLanes[4] = {4, 16, 64, 64};

// ---------------------------------------------------------------------------
// Row-block ("stream") CSR for irregular row lengths.  The rows are cut once,
// on the host, into blocks of consecutive rows holding at most RB_CAP
// entries and RB_NT rows (a longer row is a block of its own).  A work-group
// takes one block: its entries are one contiguous span of data / indices,
// read fully coalesced whatever the row lengths, and each product
// data[k] * x[indices[k]] goes to LDS; then the block's rows split the
// work-group's threads between them (L lanes per row, L the largest power of
// two with rows * L <= RB_NT) and fold their products from LDS in a fixed
// order.  Every work-group moves about the same number of entries.  A long
// row's block strides the row over all threads (fma partials, fixed tree).
constexpr int RB_NT = 256, RB_CAP = 2048;

__global__ __launch_bounds__(RB_NT) void csr_rowblock_kernel(const int *__restrict__ blocks,
                                                             const int *__restrict__ indptr,
                                                             const int *__restrict__ indices,
                                                             const float *__restrict__ data,
                                                             const float *__restrict__ x,
                                                             float *__restrict__ y) {
    __shared__ float prod[RB_CAP];
    __shared__ float red[RB_NT];
    const int tid = threadIdx.x;
    const int r0 = blocks[blockIdx.x], r1 = blocks[blockIdx.x + 1];
    const int beg = indptr[r0], end = indptr[r1];
    const int n = end - beg, nrows = r1 - r0;
    if (n > RB_CAP) {  // one long row
        float acc = 0.0f;
        for (int k = beg + tid; k < end; k += RB_NT) acc = fmaf(data[k], x[indices[k]], acc);
        red[tid] = acc;
        __syncthreads();
        for (int w = RB_NT / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        if (tid == 0) y[r0] = red[0];
        return;
    }
    for (int k = tid; k < n; k += RB_NT) prod[k] = data[beg + k] * x[indices[beg + k]];
    __syncthreads();
    int L = 1;
    while (L < 64 && 2 * L * nrows <= RB_NT) L *= 2;
    const int row = tid / L, lane = tid - row * L;  // a row's lanes sit in one wave
    float acc = 0.0f;
    if (row < nrows) {
        const int a = indptr[r0 + row] - beg, e = indptr[r0 + row + 1] - beg;
        for (int k = a + lane; k < e; k += L) acc += prod[k];
    }
    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (row < nrows && lane == 0) y[r0 + row] = acc;
}

// ---------------------------------------------------------------------------
// Column-sliced CSR for rows whose columns are spread over the whole of x
// (the power-law matrix: uniformly random columns, x = 16.8 MB, larger than
// one XCD's 4 MB L2).  The columns are cut into SL_N = 8 equal ranges, one
// per XCD, and each range's entries form their own CSR over all rows (built
// once per sparsity pattern on the host; the values are permuted into that
// order once per matrix on the device).  Work-group g serves slice g % 8,
// and work-groups are dealt to the XCDs in that order, so every gather of x
// an XCD makes falls in its own 2.1 MB slice, which stays in its L2 while the
// matrix streams past with non-temporal loads.  Each work-group takes SL_RB
// consecutive rows of its slice and writes one partial sum per row; a second
// pass adds a row's 8 partials in slice order.  Within a slice a row's
// products are added in entry order, so results are deterministic and within
// the multi-lane bound of the fp64 product.
constexpr int SL_N = 8, SL_RB = 256, SL_CAP = 2048;

__global__ __launch_bounds__(SL_RB) void csr_slice_kernel(int M, const int *__restrict__ sptr,
                                                          const int *__restrict__ sidx,
                                                          const float *__restrict__ sdata,
                                                          const float *__restrict__ x,
                                                          float *__restrict__ partial) {
    __shared__ float prod[SL_CAP];
    const int s = blockIdx.x % SL_N;
    const int r0 = (blockIdx.x / SL_N) * SL_RB;
    if (r0 >= M) return;
    const int r1 = min(r0 + SL_RB, M);
    const int *ptr = sptr + (int64_t)s * (M + 1);
    const int tid = threadIdx.x, row = r0 + tid;
    const int beg = ptr[r0], end = ptr[r1];
    int a = 0, e = 0;
    if (row < r1) {
        a = __builtin_nontemporal_load(ptr + row);
        e = __builtin_nontemporal_load(ptr + row + 1);
    }
    float acc = 0.0f;
    for (int cb = beg; cb < end; cb += SL_CAP) {
        const int m = min(SL_CAP, end - cb);
        for (int k = tid; k < m; k += SL_RB) {
            const float d = __builtin_nontemporal_load(sdata + cb + k);
            const int c = __builtin_nontemporal_load(sidx + cb + k);
            prod[k] = d * x[c];
        }
        __syncthreads();
        const int lo = max(a, cb) - cb, hi = min(e, cb + m) - cb;
        for (int k = lo; k < hi; ++k) acc += prod[k];
        __syncthreads();
    }
    if (row < r1) __builtin_nontemporal_store(acc, partial + (int64_t)s * M + row);
}

__global__ __launch_bounds__(256) void csr_slice_sum_kernel(int M, const float *__restrict__ partial,
                                                            float *__restrict__ y) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M) return;
    float acc = __builtin_nontemporal_load(partial + i);
#pragma unroll
    for (int s = 1; s < SL_N; ++s) acc += __builtin_nontemporal_load(partial + (int64_t)s * M + i);
    y[i] = acc;
}

__global__ __launch_bounds__(256) void csr_slice_values_kernel(int64_t nnz, const int *__restrict__ perm,
                                                               const float *__restrict__ data,
                                                               float *__restrict__ sdata) {
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < nnz; k += stride)
        sdata[k] = data[perm[k]];
}

}  // namespace hmx

using namespace hmx;

extern "C" int hmx_csr_slice_count(void) { return SL_N; }

extern "C" int hmx_csr_slice_build(int M, int N, const int *indptr_host, const int *indices_host,
                                   int *sptr_host, int *sidx_host, int *perm_host) {
    HMX_REQUIRE(M >= 0 && N >= 1 && indptr_host && sptr_host, "hmx_csr_slice_build: bad args");
    const int64_t nnz = indptr_host[M];
    HMX_REQUIRE(nnz == 0 || (indices_host && sidx_host && perm_host),
                "hmx_csr_slice_build: null entry arrays");
    const int sw = (N + SL_N - 1) / SL_N;
    const int64_t P = (int64_t)M + 1;
    // entries per (slice, row), then slice-major prefix sums
    std::vector<int> cnt((size_t)SL_N * M, 0);
    for (int i = 0; i < M; ++i) {
        HMX_REQUIRE(indptr_host[i + 1] >= indptr_host[i], "hmx_csr_slice_build: indptr decreases");
        for (int j = indptr_host[i]; j < indptr_host[i + 1]; ++j) {
            const int c = indices_host[j];
            HMX_REQUIRE(c >= 0 && c < N, "hmx_csr_slice_build: column index out of range");
            ++cnt[(size_t)(c / sw) * M + i];
        }
    }
    int64_t pos = 0;
    for (int sl = 0; sl < SL_N; ++sl) {
        int *ptr = sptr_host + sl * P;
        for (int i = 0; i < M; ++i) {
            ptr[i] = (int)pos;
            pos += cnt[(size_t)sl * M + i];
        }
        ptr[M] = (int)pos;
    }
    // stable fill: a row's entries keep their order inside each slice
    std::vector<int> cur((size_t)SL_N * M);
    for (int sl = 0; sl < SL_N; ++sl)
        std::copy(sptr_host + sl * P, sptr_host + sl * P + M, cur.begin() + (size_t)sl * M);
    for (int i = 0; i < M; ++i)
        for (int j = indptr_host[i]; j < indptr_host[i + 1]; ++j) {
            const int c = indices_host[j];
            const int k = cur[(size_t)(c / sw) * M + i]++;
            sidx_host[k] = c;
            perm_host[k] = j;
        }
    return HMX_OK;
}

extern "C" int hmx_csr_slice_values(int64_t nnz, const int *perm, const float *data,
                                    float *sdata, void *stream) {
    HMX_REQUIRE(nnz >= 0 && (nnz == 0 || (perm && data && sdata)),
                "hmx_csr_slice_values: bad args");
    if (nnz == 0) return HMX_OK;
    const int64_t blocks = std::min<int64_t>((nnz + 255) / 256, 16384);
    hipLaunchKernelGGL(csr_slice_values_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       as_stream(stream), nnz, perm, data, sdata);
    HMX_LAUNCH_CHECK("csr_slice_values");
    return HMX_OK;
}

extern "C" int hmx_csr_spmv_sliced(int M, const int *sptr, const int *sidx, const float *sdata,
                                   const float *x, float *partial, float *y, void *stream) {
    HMX_REQUIRE(M >= 0 && sptr && x && partial && y, "hmx_csr_spmv_sliced: null argument");
    HMX_REQUIRE((const void *)x != (const void *)y, "hmx_csr_spmv_sliced: x and y must differ");
    if (M == 0) return HMX_OK;
    hipStream_t s = as_stream(stream);
    const int64_t blocks = (int64_t)SL_N * ((M + SL_RB - 1) / SL_RB);
    HMX_REQUIRE(blocks < ((int64_t)1 << 31), "hmx_csr_spmv_sliced: too many rows");
    hipLaunchKernelGGL(csr_slice_kernel, dim3((unsigned)blocks), dim3(SL_RB), 0, s, M, sptr, sidx,
                       sdata, x, partial);
    HMX_LAUNCH_CHECK("csr_slice");
    hipLaunchKernelGGL(csr_slice_sum_kernel, dim3((M + 255) / 256), dim3(256), 0, s, M, partial,
                       y);
    HMX_LAUNCH_CHECK("csr_slice_sum");
    return HMX_OK;
}

extern "C" int hmx_csr_spmv(int M, int N, int64_t nnz, const int *indptr, const int *indices,
                            const float *data, const float *x, float *y, int lanes_per_row,
                            void *stream) {
    HMX_REQUIRE(indptr && x && y, "hmx_csr_spmv: null array");
    HMX_REQUIRE(M >= 0 && N >= 0 && nnz >= 0, "hmx_csr_spmv: negative size");
    HMX_REQUIRE(nnz < ((int64_t)1 << 31), "hmx_csr_spmv: int32 indices limit nnz to 2^31-1");
    HMX_REQUIRE(nnz == 0 || (indices && data), "hmx_csr_spmv: null indices/data");
    HMX_REQUIRE((const void *)x != (const void *)y, "hmx_csr_spmv: x and y must be distinct");
    return launch_csr(M, nnz, indptr, indices, data, x, y, lanes_per_row, as_stream(stream));
}

extern "C" int hmx_csr_bin_rows(int M, const int *indptr_host, int *bins_host, int *bin_counts) {
    HMX_REQUIRE(M >= 0 && indptr_host && bins_host && bin_counts, "hmx_csr_bin_rows: bad args");
    int cnt[4] = {0, 0, 0, 0};
    for (int i = 0; i < M; ++i) {
        const int len = indptr_host[i + 1] - indptr_host[i];
        int b = 0;
        while (len > kBinMax[b]) ++b;
        ++cnt[b];
    }
    int off[4] = {0, cnt[0], cnt[0] + cnt[1], cnt[0] + cnt[1] + cnt[2]};
    for (int i = 0; i < M; ++i) {
        const int len = indptr_host[i + 1] - indptr_host[i];
        int b = 0;
        while (len > kBinMax[b]) ++b;
        bins_host[off[b]++] = i;
    }
    for (int b = 0; b < 4; ++b) bin_counts[b] = cnt[b];
    return HMX_OK;
}

extern "C" int hmx_csr_spmv_binned(int M, const int *indptr, const int *indices,
                                   const float *data, const float *x, float *y, const int *bins,
                                   const int *bin_counts, void *stream) {
    HMX_REQUIRE(indptr && indices && data && x && y && bins && bin_counts,
                "hmx_csr_spmv_binned: null argument");
    int total = 0, off = 0;
    for (int b = 0; b < 4; ++b) total += bin_counts[b];
    HMX_REQUIRE(total == M, "hmx_csr_spmv_binned: bin counts do not add up to M");
    hipStream_t s = as_stream(stream);
    for (int b = 0; b < 4; ++b) {
        int rc = launch_csr_L(kBinLanes[b], bin_counts[b], bins + off, indptr, indices, data, x,
                              y, s);
        if (rc) return rc;
        off += bin_counts[b];
    }
    return HMX_OK;
}

extern "C" int hmx_csr_row_blocks(int M, const int *indptr_host, int *blocks_host,
                                  int *nblocks) {
    HMX_REQUIRE(M >= 0 && indptr_host && blocks_host && nblocks,
                "hmx_csr_row_blocks: bad args");
    int nb = 0, r = 0;
    blocks_host[0] = 0;
    while (r < M) {
        const int start = r;
        if (indptr_host[r + 1] - indptr_host[r] > RB_CAP) {
            ++r;  // a long row alone
        } else {
            int64_t n = 0;
            while (r < M && r - start < RB_NT &&
                   n + (indptr_host[r + 1] - indptr_host[r]) <= RB_CAP) {
                n += indptr_host[r + 1] - indptr_host[r];
                ++r;
            }
        }
        blocks_host[++nb] = r;
    }
    *nblocks = nb;
    return HMX_OK;
}

extern "C" int hmx_csr_spmv_rowblocks(int M, const int *indptr, const int *indices,
                                      const float *data, const float *x, float *y,
                                      const int *blocks, int nblocks, void *stream) {
    HMX_REQUIRE(M >= 0 && nblocks >= 0 && indptr && x && y && (blocks || nblocks == 0),
                "hmx_csr_spmv_rowblocks: null argument");
    HMX_REQUIRE((const void *)x != (const void *)y, "hmx_csr_spmv_rowblocks: x and y must differ");
    if (nblocks == 0) return HMX_OK;
    hipLaunchKernelGGL(csr_rowblock_kernel, dim3(nblocks), dim3(RB_NT), 0, as_stream(stream),
                       blocks, indptr, indices, data, x, y);
    HMX_LAUNCH_CHECK("csr_rowblock");
    return HMX_OK;
}
