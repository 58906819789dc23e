/*
 * xpu_kernels.hip - ahead-of-time kernels of libxpuhip.so: the pipeline and
 * scan + group-by kernels running the plan's programs through the op
 * interpreter (device code in xpu_device.h), and small helper kernels.
 */
#include "xpu_device.h"

template <bool PREAGG, int NR>
__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_pipeline_kernel(xpu_kparams kp)
{
    pipeline_body<PREAGG, NR, Interp>((cplan *)(uintptr_t)kp.plan, kp);
}

template <int NR, int PM>
__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_projection_kernel(xpu_kparams kp)
{
    pipeline_body<false, NR, Interp, PM>((cplan *)(uintptr_t)kp.plan, kp);
}

__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_tile_scan_kernel(uint64_t *st, uint32_t ntiles, const kern_data_store *dst, xpu_task_ctl *ctl)
{
    tile_scan_body(st, ntiles, dst, ctl);
}

template <int NR, int RPT>
__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_scan_agg_kernel(xpu_kparams kp)
{
    scan_agg_body<NR, RPT, Interp>((cplan *)(uintptr_t)kp.plan, kp);
}

template <int NR>
__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_spill_agg_kernel(xpu_kparams kp)
{
    spill_agg_body<NR, Interp>((cplan *)(uintptr_t)kp.plan, kp);
}

template <int NR>
__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_dump_merge_kernel(xpu_kparams kp)
{
    dump_merge_body<NR, Interp>((cplan *)(uintptr_t)kp.plan, kp);
}

template <int NR, int RPT>
__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_batch_agg_kernel(xpu_kparams kp)
{
    batch_agg_body<NR, RPT, Interp>((cplan *)(uintptr_t)kp.plan, kp);
}

__global__ void
xpu_batch_spill_setup_kernel(xpu_kparams kp)
{
    batch_spill_setup_body(kp);
}

template <int NR>
__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_batch_spill_kernel(xpu_kparams kp)
{
    batch_spill_body<NR, Interp>((cplan *)(uintptr_t)kp.plan, kp);
}

/* ------------------------------------------------------------------
 * helper kernels
 * ------------------------------------------------------------------ */

/* clears device memory with vector stores (no hipMemsetAsync nodes, so the
 * launch sequence stays capturable in a hipGraph) */
__global__ void
xpu_fill_kernel(uint64_t *ptr, uint64_t value, size_t nwords)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride)
        ptr[i] = value;
}

/* Bloom filter of an inner relation's join hashes: one lane per hash slot
 * walks that slot's chain */
__global__ void
xpu_bloom_build_kernel(kern_data_store *inner, uint64_t nslots, uint64_t *words, uint32_t mask,
                       uint32_t *maxdup)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < nslots; i += stride)
        bloom_build_range(inner, words, mask, i, i + 1, maxdup);
}

/* writes the kds_dst header: the session's KDS_FORMAT_ROW header with the
 * row count and usage the pipeline kernel produced */
__global__ void
xpu_finalize_dst_kernel(kern_data_store *kds_dst, const kern_data_store *head, uint32_t head_len,
                        uint64_t length, const xpu_task_ctl *ctl)
{
    for (uint32_t i = threadIdx.x * 4; i < head_len; i += blockDim.x * 4)
        *(uint32_t *)((char *)kds_dst + i) = *(const uint32_t *)((const char *)head + i);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        kds_dst->length = length;
        kds_dst->nitems = (uint32_t)ctl->out_nitems;
        kds_dst->usage = ctl->out_usage;
        kds_dst->format = KDS_FORMAT_ROW;
        kds_dst->hash_nslots = 0;
    }
}

extern "C" int
xpu_launch_pipeline(const xpu_kparams *kp, int nregs, int grid, int block, size_t lds_bytes,
                    bool preagg, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (nregs <= XP_NREGS_SMALL)
    {
        if (preagg)
            hipLaunchKernelGGL((xpu_pipeline_kernel<true, XP_NREGS_SMALL>), dim3(grid), dim3(block),
                               lds_bytes, s, *kp);
        else
            hipLaunchKernelGGL((xpu_pipeline_kernel<false, XP_NREGS_SMALL>), dim3(grid), dim3(block),
                               lds_bytes, s, *kp);
    }
    else
    {
        if (preagg)
            hipLaunchKernelGGL((xpu_pipeline_kernel<true, XP_NREGS>), dim3(grid), dim3(block),
                               lds_bytes, s, *kp);
        else
            hipLaunchKernelGGL((xpu_pipeline_kernel<false, XP_NREGS>), dim3(grid), dim3(block),
                               lds_bytes, s, *kp);
    }
    return (int)hipGetLastError();
}

template <int NR>
static void
launch_scan_agg(const xpu_kparams *kp, int rpt, int grid, int block, size_t lds, hipStream_t s)
{
    if (rpt >= 4)
        hipLaunchKernelGGL((xpu_scan_agg_kernel<NR, 4>), dim3(grid), dim3(block), lds, s, *kp);
    else if (rpt == 2)
        hipLaunchKernelGGL((xpu_scan_agg_kernel<NR, 2>), dim3(grid), dim3(block), lds, s, *kp);
    else
        hipLaunchKernelGGL((xpu_scan_agg_kernel<NR, 1>), dim3(grid), dim3(block), lds, s, *kp);
}

extern "C" int
xpu_launch_scan_agg(const xpu_kparams *kp, int nregs, int rpt, int grid, int block,
                    size_t lds_bytes, void *stream)
{
    if (nregs <= XP_NREGS_SMALL)
        launch_scan_agg<XP_NREGS_SMALL>(kp, rpt, grid, block, lds_bytes, (hipStream_t)stream);
    else
        launch_scan_agg<XP_NREGS>(kp, rpt, grid, block, lds_bytes, (hipStream_t)stream);
    return (int)hipGetLastError();
}

template <int NR>
static void
launch_batch(const xpu_kparams *kp, int rpt, int grid, int block, size_t lds, hipStream_t s)
{
    if (rpt >= 4)
        hipLaunchKernelGGL((xpu_batch_agg_kernel<NR, 4>), dim3(grid), dim3(block), lds, s, *kp);
    else if (rpt == 2)
        hipLaunchKernelGGL((xpu_batch_agg_kernel<NR, 2>), dim3(grid), dim3(block), lds, s, *kp);
    else
        hipLaunchKernelGGL((xpu_batch_agg_kernel<NR, 1>), dim3(grid), dim3(block), lds, s, *kp);
    hipLaunchKernelGGL(xpu_batch_spill_setup_kernel, dim3(1), dim3(64), 0, s, *kp);
    hipLaunchKernelGGL((xpu_batch_spill_kernel<NR>), dim3(grid), dim3(block), lds, s, *kp);
}

extern "C" int
xpu_launch_batch_setup(const xpu_kparams *kp, void *stream)
{
    hipLaunchKernelGGL(xpu_batch_spill_setup_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, *kp);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_batch(const xpu_kparams *kp, int nregs, int rpt, int grid, int block,
                 size_t lds_bytes, void *stream)
{
    if (nregs <= XP_NREGS_SMALL)
        launch_batch<XP_NREGS_SMALL>(kp, rpt, grid, block, lds_bytes, (hipStream_t)stream);
    else
        launch_batch<XP_NREGS>(kp, rpt, grid, block, lds_bytes, (hipStream_t)stream);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_dump_merge(const xpu_kparams *kp, int nregs, int grid, int block, size_t lds_bytes,
                      void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (nregs <= XP_NREGS_SMALL)
        hipLaunchKernelGGL((xpu_dump_merge_kernel<XP_NREGS_SMALL>), dim3(grid), dim3(block),
                           lds_bytes, s, *kp);
    else
        hipLaunchKernelGGL((xpu_dump_merge_kernel<XP_NREGS>), dim3(grid), dim3(block),
                           lds_bytes, s, *kp);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_spill_agg(const xpu_kparams *kp, int nregs, int grid, int block, size_t lds_bytes,
                     void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (nregs <= XP_NREGS_SMALL)
        hipLaunchKernelGGL((xpu_spill_agg_kernel<XP_NREGS_SMALL>), dim3(grid), dim3(block),
                           lds_bytes, s, *kp);
    else
        hipLaunchKernelGGL((xpu_spill_agg_kernel<XP_NREGS>), dim3(grid), dim3(block),
                           lds_bytes, s, *kp);
    return (int)hipGetLastError();
}

template <int NR>
__global__ void __launch_bounds__(256)
xpu_final_merge_ovf_kernel(const xplan *plan, kern_data_store *kds_final,
                           const kern_data_store *kds_ovf, xpu_task_ctl *ctl)
{
    final_merge_ovf_body<NR>((cplan *)(uintptr_t)plan, kds_final, kds_ovf, ctl);
}

extern "C" int
xpu_launch_final_merge_ovf(const xplan *plan, int nregs, kern_data_store *kds_final,
                           const kern_data_store *kds_ovf, uint32_t ovf_nitems, xpu_task_ctl *ctl,
                           void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const uint32_t block = 256;
    uint32_t grid = (ovf_nitems + block - 1) / block;
    if (grid == 0)
        return 0;
    if (grid > 4096)
        grid = 4096;
    if (nregs <= XP_NREGS_SMALL)
        hipLaunchKernelGGL((xpu_final_merge_ovf_kernel<XP_NREGS_SMALL>), dim3(grid), dim3(block),
                           0, s, plan, kds_final, kds_ovf, ctl);
    else
        hipLaunchKernelGGL((xpu_final_merge_ovf_kernel<XP_NREGS>), dim3(grid), dim3(block),
                           0, s, plan, kds_final, kds_ovf, ctl);
    return (int)hipGetLastError();
}

/* device-to-device copy in 16-byte words (plus a byte tail) */
__global__ void
xpu_copy_kernel(uint4 *dst, const uint4 *src, size_t nwords, char *dtail, const char *stail,
                size_t ntail)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < nwords; i += stride)
        dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < ntail)
        dtail[threadIdx.x] = stail[threadIdx.x];
}

extern "C" int
xpu_launch_copy(void *dst, const void *src, size_t nbytes, void *stream)
{
    if (nbytes == 0)
        return 0;
    size_t nwords = nbytes / 16;
    size_t ntail = nbytes % 16;
    size_t blocks = (nwords + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    if (blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL(xpu_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (uint4 *)dst, (const uint4 *)src, nwords, (char *)dst + nwords * 16,
                       (const char *)src + nwords * 16, ntail);
    return (int)hipGetLastError();
}

__global__ void __launch_bounds__(XP_BLOCK_MAX)
xpu_tile_scan_blocks_kernel(uint64_t *st, const uint64_t *bt, uint32_t nwt, const kern_data_store *dst,
                            xpu_task_ctl *ctl)
{
    tile_scan_blocks_body(st, bt, nwt, dst, ctl);
}

extern "C" int
xpu_launch_tile_scan_blocks(const xpu_kparams *kp, int grid, void *stream)
{
    hipLaunchKernelGGL(xpu_tile_scan_blocks_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                       kp->tile_status, kp->block_status, kp->ntiles, (const kern_data_store *)kp->kds_out,
                       kp->ctl);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_tile_scan(const xpu_kparams *kp, void *stream)
{
    hipLaunchKernelGGL(xpu_tile_scan_kernel, dim3(1), dim3(XP_BLOCK_MAX), 0, (hipStream_t)stream,
                       kp->tile_status, kp->ntiles, (const kern_data_store *)kp->kds_out, kp->ctl);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_projection(const xpu_kparams *kp, int nregs, int grid, int block, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (nregs <= XP_NREGS_SMALL)
        hipLaunchKernelGGL((xpu_projection_kernel<XP_NREGS_SMALL, PM_COUNT>), dim3(grid), dim3(block),
                           0, s, *kp);
    else
        hipLaunchKernelGGL((xpu_projection_kernel<XP_NREGS, PM_COUNT>), dim3(grid), dim3(block),
                           0, s, *kp);
    int rc = xpu_launch_tile_scan(kp, stream);
    if (rc)
        return rc;
    if (nregs <= XP_NREGS_SMALL)
        hipLaunchKernelGGL((xpu_projection_kernel<XP_NREGS_SMALL, PM_WRITE>), dim3(grid), dim3(block),
                           0, s, *kp);
    else
        hipLaunchKernelGGL((xpu_projection_kernel<XP_NREGS, PM_WRITE>), dim3(grid), dim3(block),
                           0, s, *kp);
    return (int)hipGetLastError();
}

/* ------------------------------------------------------------------
 * pinned inner buffers: KDS_FORMAT_HASH chunks kept in device memory
 * ------------------------------------------------------------------ */

/* One thread: reserves the rows and bytes of the ROW chunk rk (rows as
 * kern_hashitems) in hk under the buffer's lock, nitems = UINT_MAX while
 * held (the allocation protocol of cuda_gpujoin.cu:836-878).  base[0] and
 * base[1] get the first row id and byte offset, or base[0] = ~0 and base[1]
 * = 1 (the buffer is full) or 2 (the lock was not obtained). */
__global__ void
xpu_hash_reserve_kernel(kern_data_store *hk, const kern_data_store *rk, uint64_t *base)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    const uint64_t n = rk->nitems;
    const uint64_t bytes = rk->usage + 8 * n;
    for (uint32_t spin = 0; spin < XLB_SPIN_LIMIT; spin++)
    {
        uint32_t cur = __hip_atomic_load(&hk->nitems, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == 0xffffffffU || atomicCAS(&hk->nitems, cur, 0xffffffffU) != cur)
            continue;
        const uint64_t usage = __hip_atomic_load(&hk->usage, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t need = KDS_HEAD_LENGTH(hk) + 8 * (hk->hash_nslots + cur + n) + usage + bytes;
        if (cur + n >= 0xffffffffULL || need > hk->length)
        {
            base[0] = ~0ULL;
            base[1] = 1;
        }
        else
        {
            __hip_atomic_store(&hk->usage, usage + bytes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            base[0] = cur;
            base[1] = usage;
            cur += (uint32_t)n;
        }
        __hip_atomic_store(&hk->nitems, cur, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    base[0] = ~0ULL;
    base[1] = 2;
}

/* Appends every row of the ROW chunk rk (rows packed from its tail in row
 * order) into hk at the reserved base: row i becomes a kern_hashitem
 * base[1] + U_i + 8 (i + 1) bytes from hk's tail (U_i: its tuple's offset in
 * rk), takes row id base[0] + i and is pushed onto the chain of its
 * tupitem's hash (the HASH branch of cuda_gpujoin.cu:898-916). */
__global__ void
xpu_hash_append_kernel(kern_data_store *hk, const kern_data_store *rk, const uint64_t *base)
{
    const uint64_t b0 = base[0], u0 = base[1];
    if (b0 == ~0ULL)
        return;
    const uint32_t n = rk->nitems;
    const uint64_t *rrow = (const uint64_t *)((const char *)rk + KDS_HEAD_LENGTH(rk));
    unsigned long long *slots = (unsigned long long *)((char *)hk + KDS_HEAD_LENGTH(hk));
    const uint64_t nslots = hk->hash_nslots;
    uint64_t *hrow = (uint64_t *)(slots + nslots);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint64_t U = rrow[i];
        if (U > rk->usage || U < 16)
            continue;
        const kern_tupitem *t = (const kern_tupitem *)((const char *)rk + rk->length - U);
        const uint32_t t_len = t->t_len;
        if (t_len > U || t_len < KERN_TUPITEM_T_BITS_OFFSET + ROWID_SIZE)
            continue;
        const uint64_t H = u0 + U + 8ULL * (i + 1);
        kern_hashitem *hi = (kern_hashitem *)((char *)hk + hk->length - H);
        const uint64_t *src = (const uint64_t *)t;
        uint64_t *dst = (uint64_t *)&hi->t;
        for (uint32_t w = 0; w < (t_len + 7) / 8; w++)
            dst[w] = src[w];
        *(uint32_t *)((char *)&hi->t + t_len - ROWID_SIZE) = (uint32_t)(b0 + i);
        hi->next = atomicExch(&slots[t->hash % nslots], (unsigned long long)H);
        hrow[b0 + i] = H - offsetof(kern_hashitem, t);
    }
}

/* Re-chains the items of a KDS_FORMAT_HASH chunk whose row index and tuple
 * area were copied from another buffer, into its own (zeroed) hash slots:
 * kern_buffer_reconstruction (cuda_gpujoin.cu:18-74) with the items left
 * where they are. */
__global__ void
xpu_hash_rechain_kernel(kern_data_store *hk)
{
    const uint32_t n = hk->nitems;
    unsigned long long *slots = (unsigned long long *)((char *)hk + KDS_HEAD_LENGTH(hk));
    const uint64_t nslots = hk->hash_nslots;
    const uint64_t *hrow = (const uint64_t *)(slots + nslots);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    {
        const uint64_t H = hrow[i] + offsetof(kern_hashitem, t);
        if (H > hk->usage)
            continue;
        kern_hashitem *hi = (kern_hashitem *)((char *)hk + hk->length - H);
        hi->next = atomicExch(&slots[hi->t.hash % nslots], (unsigned long long)H);
    }
}

extern "C" int
xpu_launch_hash_append(kern_data_store *hk, const kern_data_store *rk, uint64_t *base,
                       uint32_t nrows, void *stream)
{
    hipLaunchKernelGGL(xpu_hash_reserve_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream,
                       hk, rk, base);
    if (nrows > 0)
    {
        uint32_t blocks = (nrows + 255) / 256;
        if (blocks > 4096)
            blocks = 4096;
        hipLaunchKernelGGL(xpu_hash_append_kernel, dim3(blocks), dim3(256), 0,
                           (hipStream_t)stream, hk, rk, (const uint64_t *)base);
    }
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_hash_rechain(kern_data_store *hk, uint32_t nitems, void *stream)
{
    if (nitems == 0)
        return 0;
    uint32_t blocks = (nitems + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(xpu_hash_rechain_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, hk);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_fill(void *ptr, uint64_t value, size_t nwords, void *stream)
{
    if (nwords == 0)
        return 0;
    size_t blocks = (nwords + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(xpu_fill_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, (uint64_t *)ptr, value, nwords);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_bloom_build(kern_data_store *inner, uint64_t nslots, uint64_t *words, uint32_t mask,
                       uint32_t *maxdup, void *stream)
{
    if (nslots == 0)
        return 0;
    uint64_t blocks = (nslots + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(xpu_bloom_build_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, inner, nslots, words, mask, maxdup);
    return (int)hipGetLastError();
}

__global__ void
xpu_bloom_fold_kernel(const uint64_t *words, uint64_t nwords, uint64_t *folded, uint64_t fwords)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fwords)
        return;
    uint64_t w = 0;
    for (uint64_t j = i; j < nwords; j += fwords)
        w |= words[j];
    folded[i] = w;
}

extern "C" int
xpu_launch_bloom_fold(const uint64_t *words, uint64_t nwords, uint64_t *folded, uint64_t fwords,
                      void *stream)
{
    if (fwords == 0)
        return 0;
    hipLaunchKernelGGL(xpu_bloom_fold_kernel, dim3((unsigned)((fwords + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, words, nwords, folded, fwords);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_finalize_dst(kern_data_store *kds_dst, const kern_data_store *head, uint32_t head_len,
                        uint64_t length, const xpu_task_ctl *ctl, void *stream)
{
    hipLaunchKernelGGL(xpu_finalize_dst_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream,
                       kds_dst, head, head_len, length, ctl);
    return (int)hipGetLastError();
}
