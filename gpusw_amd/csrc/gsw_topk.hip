// Device top-k over the rows of an int32 score matrix.
//
// Order: score descending, ties to the lower column index -- the order of a stable
// descending sort of each row (the reference's dense AlignResult.top_k,
// src/gpusw/result.py:30-47).  One workgroup per row: every thread keeps the KL best
// entries of its strided share of the row; k rounds each pick the workgroup-best head
// and pop it.  A thread whose list runs dry refills it from its share, keeping only
// entries ranked after the last one it gave up, so any k <= ncols is exact.
#include <algorithm>
#include <chrono>
#include <climits>

#include "gsw_batch.h"
#include "gsw_internal.h"

namespace gsw {

constexpr int TOPK_THREADS = 256;
constexpr int TOPK_KL = 8;

// Rank key: larger key = better (higher score, then lower column index).
__device__ __forceinline__ uint64_t rank_key(int score, int col) {
    return ((uint64_t)((uint32_t)score ^ 0x80000000u) << 32) | (uint32_t)(0x7fffffff - col);
}

__global__ void __launch_bounds__(TOPK_THREADS)
topk_rows_kernel(const int32_t *__restrict__ scores, int32_t ncols, int32_t k,
                 int32_t *__restrict__ out_idx, int32_t *__restrict__ out_sc) {
    __shared__ uint64_t s_red[TOPK_THREADS / 64];
    const int64_t row = blockIdx.x;
    const int32_t *srow = scores + row * (int64_t)ncols;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    uint64_t list[TOPK_KL];
    int cnt = 0;
    uint64_t floor_key = ~0ull;   // only entries ranked strictly below this are taken
    bool exhausted = false;

    auto refill = [&]() {
        cnt = 0;
        for (int j = tid; j < ncols; j += TOPK_THREADS) {
            const uint64_t key = rank_key(srow[j], j);
            if (key >= floor_key) continue;
            if (cnt < TOPK_KL) {
                int p = cnt++;
                while (p > 0 && list[p - 1] < key) { list[p] = list[p - 1]; --p; }
                list[p] = key;
            } else if (key > list[TOPK_KL - 1]) {
                int p = TOPK_KL - 1;
                while (p > 0 && list[p - 1] < key) { list[p] = list[p - 1]; --p; }
                list[p] = key;
            }
        }
        exhausted = cnt == 0;
    };
    refill();

    for (int r = 0; r < k; ++r) {
        if (cnt == 0 && !exhausted) refill();
        const uint64_t head = cnt > 0 ? list[0] : 0ull;
        uint64_t m = head;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const uint64_t o = ((uint64_t)__shfl_xor((unsigned)(m >> 32), off, 64) << 32) |
                               (uint64_t)__shfl_xor((unsigned)m, off, 64);
            m = o > m ? o : m;
        }
        if (lane == 0) s_red[wave] = m;
        __syncthreads();
        uint64_t best = s_red[0];
        for (int w = 1; w < TOPK_THREADS / 64; ++w) best = s_red[w] > best ? s_red[w] : best;
        __syncthreads();
        if (tid == 0) {
            out_sc[row * k + r] = (int32_t)((uint32_t)(best >> 32) ^ 0x80000000u);
            out_idx[row * k + r] = 0x7fffffff - (int32_t)(uint32_t)best;
        }
        if (cnt > 0 && head == best) {   // keys are unique per column: one owner
            for (int p = 1; p < cnt; ++p) list[p - 1] = list[p];
            --cnt;
            floor_key = best;
        }
    }
}

void topk_launch(const int32_t *d_scores, int32_t nrows, int32_t ncols, int32_t k, int32_t *d_idx,
                 int32_t *d_sc, hipStream_t stream) {
    if (nrows <= 0 || k <= 0) return;
    hipLaunchKernelGGL(topk_rows_kernel, dim3(nrows), dim3(TOPK_THREADS), 0, stream, d_scores, ncols, k, d_idx,
                       d_sc);
}

}  // namespace gsw

using namespace gsw;

extern "C" int gsw_topk_rows(const int32_t *d_scores, int32_t nrows, int32_t ncols, int32_t k,
                             int32_t *h_idx, int32_t *h_sc, gsw_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (nrows < 0 || ncols < 0 || k < 0) return fail(GSW_ERR_INVALID, "negative sizes");
    if (k > ncols) return fail(GSW_ERR_INVALID, "k=%d exceeds the %d columns", k, ncols);
    DeviceCtx *ctx = nullptr;
    GSW_TRY(ctx_get(&ctx));
    std::lock_guard<std::mutex> lk(ctx->mu);
    float ms = 0.f;
    if (nrows > 0 && k > 0) {
        const size_t bytes = sizeof(int32_t) * (size_t)nrows * k;
        GSW_TRY(ensure_buffer(&ctx->meta, &ctx->meta_bytes, 2 * bytes + 256));
        int32_t *d_idx = static_cast<int32_t *>(ctx->meta);
        int32_t *d_sc = d_idx + (size_t)nrows * k;
        GSW_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        topk_launch(d_scores, nrows, ncols, k, d_idx, d_sc, ctx->stream);
        GSW_HIP_TRY(hipGetLastError());
        GSW_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        GSW_HIP_TRY(hipMemcpyAsync(h_idx, d_idx, bytes, hipMemcpyDeviceToHost, ctx->stream));
        GSW_HIP_TRY(hipMemcpyAsync(h_sc, d_sc, bytes, hipMemcpyDeviceToHost, ctx->stream));
        GSW_HIP_TRY(hipStreamSynchronize(ctx->stream));
        GSW_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    }
    if (stats) {
        *stats = gsw_stats{};
        stats->kernel_ms = ms;
        stats->launches = 1;
        stats->variant = -1;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return GSW_OK;
}
