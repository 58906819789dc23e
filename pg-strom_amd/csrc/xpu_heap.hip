/*
 * xpu_heap.hip - heap-page (KDS_FORMAT_BLOCK) and row (KDS_FORMAT_ROW/HASH)
 * sources, read into column tiles for the pipeline.
 *
 * The reference feeds such chunks to its scan loop tuple by tuple: one warp
 * per page walks the line pointers into an LDS ring (cuda_gpuscan.cu:405-534)
 * and every tuple is deformed by ExecLoadVarsOuterHeap (xpu_common.cu:60-130,
 * 194-257, 1587-1620); row chunks are read per row after an MVCC check of
 * the embedded transaction attributes (cuda_gpuscan.cu:610-700).
 *
 * Here the chunk is converted once, at HBM speed, into the columnar layout
 * the scan, join and group-by kernels already read:
 *   count pass   one wave per page (or per 64 row-index entries): the
 *                tuples that the reference would load (LP_NORMAL items /
 *                visible rows), validated against the page and chunk bounds
 *   unit scan    exclusive prefix of the per-unit counts: every page owns a
 *                contiguous range of output rows, so rows keep page order
 *   deform pass  persistent waves, one page at a time each: the 8 KB page
 *                is copied into the wave's LDS slot with coalesced 8-byte
 *                loads (the next page's loads overlap this page's work),
 *                the normal items are
 *                compacted into a list, and each lane deforms one tuple from
 *                LDS (attribute walk with the reference's alignment rules),
 *                writing fixed-width values, validity bits and (offset,
 *                length) pairs of text datums for the referenced attributes
 *                only.  Rows of a wave are consecutive, so value stores are
 *                coalesced and validity bits leave as ballots.
 */
#include <hip/hip_runtime.h>

#include "xpu_kernels.h"

/* walks of up to this many attributes keep the column table in registers */
#define XHEAP_REG_COLS  8

namespace {

__device__ __forceinline__ uint32_t
lane_id()
{
    return __lane_id();
}

__device__ __forceinline__ void
heap_error(xheap_status *st, uint32_t msgid)
{
    if (atomicCAS(&st->errcode, 0U, (uint32_t)ERRCODE_DEVICE_ERROR) == 0U)
        __hip_atomic_store(&st->errmsg, msgid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* line pointers of a page: PageGetMaxOffsetNumber (xpu_common.h:1123-1127);
 * a pd_lower beyond the page is reported and reads as no items */
__device__ __forceinline__ uint32_t
page_nlp(uint32_t pd_lower, bool *bad)
{
    if (pd_lower > XPU_BLCKSZ)
    {
        *bad = true;
        return 0;
    }
    return pd_lower <= XPU_PAGE_HEADER_SIZE ? 0 : (pd_lower - XPU_PAGE_HEADER_SIZE) / 4;
}

/* an item the scan loads: ItemIdIsNormal, with the tuple inside the page */
__device__ __forceinline__ bool
page_item_ok(uint32_t lp, bool *bad)
{
    if (XPU_ITEMID_FLAGS(lp) != LP_NORMAL)
        return false;
    uint32_t off = XPU_ITEMID_OFF(lp), len = XPU_ITEMID_LEN(lp);
    if (off < XPU_PAGE_HEADER_SIZE || len < XPU_SIZEOF_HTUP_HEADER || off + len > XPU_BLCKSZ)
    {
        *bad = true;
        return false;
    }
    return true;
}

/* the row-index entry `row` of a ROW/HASH chunk: false for an empty entry or
 * a tuple that is not visible (__gpuscan_check_row_visibility,
 * cuda_gpuscan.cu:614-642).  On success *hoff is the offset from the chunk
 * head of the tuple's virtual HeapTupleHeaderData (the kern_tupitem starts
 * MINIMAL_TUPLE_OFFSET bytes into it) and *hend the end of its item. */
__device__ __forceinline__ bool
row_item(const xheap_params &p, uint32_t row, uint64_t *hoff, uint32_t *hend, bool *bad)
{
    const uint8_t *base = (const uint8_t *)p.src;
    uint64_t off = ((const uint64_t *)(base + p.rowindex_off))[row];
    if (off == 0)
        return false;
    uint64_t body_end = p.rowindex_off + 8ULL * p.nitems;
    if (off < 16 || off > p.src_length - body_end || ((p.src_length - off) & 7) != 0)
    {
        *bad = true;
        return false;
    }
    uint64_t tpos = p.src_length - off;
    const uint8_t *t = base + tpos;
    uint32_t t_len = *(const uint32_t *)t;
    uint32_t t_hoff = t[14];
    if (t_len < KERN_TUPITEM_T_BITS_OFFSET || t_len > off ||
        t_hoff < MINIMAL_TUPLE_OFFSET + KERN_TUPITEM_T_BITS_OFFSET ||
        t_hoff - MINIMAL_TUPLE_OFFSET > t_len)
    {
        *bad = true;
        return false;
    }
    if (t[8])
    {
        /* kern_tupitem_xact_attrs right in front of the payload */
        if (t_hoff < MINIMAL_TUPLE_OFFSET + KERN_TUPITEM_T_BITS_OFFSET +
                     sizeof(kern_tupitem_xact_attrs) || (t_hoff & 3) != 0)
        {
            *bad = true;
            return false;
        }
        const uint32_t *xa = (const uint32_t *)(t + t_hoff - MINIMAL_TUPLE_OFFSET -
                                                sizeof(kern_tupitem_xact_attrs));
        uint32_t xmin = xa[0], xmax = xa[1];
        if (xmin == InvalidTransactionId)
            return false;
        if (xmin != FrozenTransactionId)
        {
            bool found = false;
            for (uint32_t i = 0; i < p.nxids; i++)
                found |= (p.xids[i] == xmin);
            if (!found)
                return false;
        }
        if (xmax != InvalidTransactionId)
        {
            if (xmax == FrozenTransactionId)
                return false;
            for (uint32_t i = 0; i < p.nxids; i++)
                if (p.xids[i] == xmax)
                    return false;
        }
    }
    *hoff = tpos - MINIMAL_TUPLE_OFFSET;
    *hend = MINIMAL_TUPLE_OFFSET + t_len;
    return true;
}

/* little-endian value of w bytes at q (w in 1, 2, 4, 8) */
__device__ __forceinline__ uint64_t
load_le(const uint8_t *q, uint32_t w, bool aligned)
{
    if (aligned)
    {
        switch (w)
        {
            case 1: return q[0];
            case 2: return *(const uint16_t *)q;
            case 4: return *(const uint32_t *)q;
            default: return *(const uint64_t *)q;
        }
    }
    uint64_t v = 0;
    for (uint32_t i = 0; i < w; i++)
        v |= (uint64_t)q[i] << (8 * i);
    return v;
}

/* bits of 64 consecutive rows starting at row0 into a word bitmap */
__device__ __forceinline__ void
put_bits(uint32_t *words, uint32_t row0, uint64_t mask)
{
    uint32_t s = row0 & 31, l = lane_id();
    if (l < 3 && mask)
    {
        uint32_t v;
        if (s == 0)
            v = l == 0 ? (uint32_t)mask : l == 1 ? (uint32_t)(mask >> 32) : 0U;
        else
            v = l == 0 ? (uint32_t)(mask << s)
              : l == 1 ? (uint32_t)(mask >> (32 - s))
                       : (uint32_t)(mask >> (64 - s));
        if (v)
            atomicOr(&words[(row0 >> 5) + l], v);
    }
}

/*
 * Deforms the tuple whose HeapTupleHeaderData is at h + t (its item ends at
 * h + end) into row `row` of the tile buffer; lanes 0..n-1 of the wave are
 * active and hold rows row0..row0+n-1.  The walk follows
 * __kern_extract_common_tuple (xpu_common.cu:60-130): NULL attributes take
 * no space, fixed-width ones are aligned to attalign, varlena ones are
 * aligned unless their first byte is non-zero (a short header), sizes by
 * VARSIZE_ANY; positions are aligned relative to the payload.  h_src is the
 * offset of h from the chunk head (text datums are addressed by it).
 * NC > 0: the column table (at most NC entries) is in registers, cols[j]
 * with the walk unrolled; NC == 0: cols is the table in memory.
 */
template <int NC>
__device__ __forceinline__ void
deform_tuple(const xheap_params &p, const xheap_col *cols, const uint8_t *h, uint32_t t,
             uint32_t end, uint64_t h_src, uint32_t row, uint32_t row0, bool active, bool *bad)
{
    uint32_t natts = 0, hasnull = 0, p0 = 0;
    if (active)
    {
        uint32_t im2 = (uint32_t)h[t + XPU_HTUP_INFOMASK2] | ((uint32_t)h[t + XPU_HTUP_INFOMASK2 + 1] << 8);
        uint32_t im = h[t + XPU_HTUP_INFOMASK];
        uint32_t hoff = h[t + XPU_HTUP_HOFF];
        natts = im2 & HEAP_NATTS_MASK;
        if (natts > p.src_ncols)
            natts = p.src_ncols;
        hasnull = im & HEAP_HASNULL;
        p0 = t + hoff;
        if (hoff < XPU_SIZEOF_HTUP_HEADER || p0 > end ||
            (hasnull && t + XPU_HTUP_BITS + (natts + 7) / 8 > p0))
        {
            *bad = true;
            active = false;
        }
    }
    uint32_t pos = p0;
    char *dst = p.dst;
    const uint32_t nwalk = NC > 0 ? (p.natts_walk < (uint32_t)NC ? p.natts_walk : (uint32_t)NC)
                                  : p.natts_walk;
#pragma unroll
    for (uint32_t j = 0; j < (NC > 0 ? (uint32_t)NC : nwalk); j++)
    {
        if (NC > 0 && j >= nwalk)
            break;
        const xheap_col c = cols[j];
        bool isnull = !active || j >= natts ||
                      (hasnull && !((h[t + XPU_HTUP_BITS + (j >> 3)] >> (j & 7)) & 1));
        uint64_t v = 0;
        uint32_t voff = 0, vlen = 0;
        if (!isnull)
        {
            if (c.attlen > 0)
            {
                pos = p0 + (uint32_t)XPU_TYPEALIGN(c.attalign, pos - p0);
                if (pos + (uint32_t)c.attlen > end)
                {
                    *bad = true;
                    isnull = true;
                    active = false;
                }
                else
                {
                    if (c.kind == XH_FIXED || c.kind == XH_BOOL)
                        v = load_le(h + pos, c.width, (pos & (c.width - 1)) == 0);
                    pos += (uint32_t)c.attlen;
                }
            }
            else
            {
                uint32_t b0 = pos < end ? h[pos] : 0xffU;
                if (b0 == 0)
                {
                    pos = p0 + (uint32_t)XPU_TYPEALIGN(c.attalign, pos - p0);
                    b0 = pos < end ? h[pos] : 0xffU;
                }
                uint32_t sz = 0, hdr = 0;
                bool ext = false;
                if (pos >= end)
                    sz = 0xffffffffU;
                else if (b0 == 0x01)
                {
                    sz = pos + 1 < end ? XPU_VARSIZE_EXTERNAL(h[pos + 1]) : 0xffffffffU;
                    ext = true;
                }
                else if (b0 & 0x01)
                {
                    sz = b0 >> 1;
                    hdr = 1;
                }
                else if (pos + 4 <= end)
                {
                    uint32_t w = (uint32_t)load_le(h + pos, 4, (pos & 3) == 0);
                    sz = (w >> 2) & 0x3fffffffU;
                    hdr = 4;
                    ext = (b0 & 0x03) == 0x02;
                    if (sz < 4)
                        sz = 0xffffffffU;
                }
                else
                    sz = 0xffffffffU;
                if (sz == 0xffffffffU || sz > end - pos)
                {
                    *bad = true;
                    isnull = true;
                    active = false;
                }
                else
                {
                    /* a compressed or external datum is addressed with its
                     * header: it is copied as it is (projection) or falls
                     * back to the CPU (anything that reads its bytes) */
                    voff = (uint32_t)(h_src + pos + (ext ? 0 : hdr));
                    vlen = ext ? XHEAP_LEN_EXTERNAL : sz - hdr;
                    pos += sz;
                }
            }
        }
        if (c.kind == XH_SKIP)
            continue;
        uint64_t valid = __ballot(active && !isnull);
        put_bits((uint32_t *)(dst + c.nullmap_off), row0, valid);
        if (c.kind == XH_BOOL)
        {
            uint64_t bits = __ballot(active && !isnull && v != 0);
            put_bits((uint32_t *)(dst + c.values_off), row0, bits);
        }
        else if (active)
        {
            char *q = dst + c.values_off;
            if (c.kind == XH_VAR)
                ((uint2 *)q)[row] = make_uint2(voff, vlen);
            else
            {
                switch (c.width)
                {
                    case 1: ((uint8_t *)q)[row] = (uint8_t)v; break;
                    case 2: ((uint16_t *)q)[row] = (uint16_t)v; break;
                    case 4: ((uint32_t *)q)[row] = (uint32_t)v; break;
                    default: ((uint64_t *)q)[row] = v; break;
                }
            }
        }
    }
}

}   /* namespace */

/* per-page (BLOCK) or per-64-row (ROW/HASH) counts of the tuples to load */
__global__ void __launch_bounds__(XHEAP_COUNT_BLOCK)
xpu_heap_count_kernel(xheap_params p)
{
    const uint32_t unit = blockIdx.x * (XHEAP_COUNT_BLOCK / 64) + threadIdx.x / 64;
    const uint32_t lane = lane_id();
    if (unit >= p.nunits)
        return;
    bool bad = false;
    uint32_t cnt = 0;
    if (p.format == KDS_FORMAT_BLOCK)
    {
        const uint8_t *pg = (const uint8_t *)p.src + p.block_offset + (uint64_t)unit * XPU_BLCKSZ;
        uint32_t nlp = page_nlp(*(const uint16_t *)(pg + XPU_PAGE_PD_LOWER), &bad);
        const uint32_t *lps = (const uint32_t *)(pg + XPU_PAGE_HEADER_SIZE);
        for (uint32_t k0 = 0; k0 < nlp; k0 += 64)
        {
            uint32_t k = k0 + lane;
            bool ok = k < nlp && page_item_ok(lps[k], &bad);
            cnt += (uint32_t)__popcll(__ballot(ok));
        }
    }
    else
    {
        uint32_t row = unit * 64 + lane;
        uint64_t hoff;
        uint32_t hend;
        bool ok = row < p.nitems && row_item(p, row, &hoff, &hend, &bad);
        cnt = (uint32_t)__popcll(__ballot(ok));
    }
    if (__ballot(bad))
    {
        if (lane == 0)
            heap_error(p.st, XMSG_HEAP_ITEM);
    }
    if (lane == 0)
        p.units[unit] = cnt;
}

/* exclusive prefix of units[0..n) in place; units[n] and st->total get the
 * sum.  Rounds of 4 * XHEAP_SCAN_BLOCK units: every thread holds 4
 * contiguous counts (one coalesced 16-byte load; the next round's load is
 * issued before this round's scan), the block scans the thread sums, and
 * the round's total carries into the next round.  The units array is padded
 * to 4 entries past n, which read as 0 and receive the total. */
__global__ void __launch_bounds__(XHEAP_SCAN_BLOCK)
xpu_heap_scan_kernel(uint32_t *units, uint32_t n, xheap_status *st)
{
    constexpr uint32_t ROUND = 4 * XHEAP_SCAN_BLOCK;
    __shared__ uint32_t wsum[XHEAP_SCAN_BLOCK / 64];
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    uint4 *u4 = (uint4 *)units;
    uint32_t carry = 0;
    uint4 cur = make_uint4(0, 0, 0, 0);
    if (4 * tid < n)
        cur = u4[tid];
    for (uint32_t r0 = 0; r0 < n; r0 += ROUND)
    {
        const uint32_t b = r0 + 4 * tid;
        uint4 nxt = make_uint4(0, 0, 0, 0);
        if (b + ROUND < n)
            nxt = u4[(b + ROUND) / 4];
        uint32_t c0 = b < n ? cur.x : 0, c1 = b + 1 < n ? cur.y : 0;
        uint32_t c2 = b + 2 < n ? cur.z : 0, c3 = b + 3 < n ? cur.w : 0;
        uint32_t s = c0 + c1 + c2 + c3;
        uint32_t incl = s;
        for (uint32_t off = 1; off < 64; off <<= 1)
        {
            uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= off)
                incl += v;
        }
        __syncthreads();
        if (lane == 63)
            wsum[wave] = incl;
        __syncthreads();
        uint32_t before = carry, round = 0;
        for (uint32_t w = 0; w < XHEAP_SCAN_BLOCK / 64; w++)
        {
            uint32_t ws = wsum[w];
            before += w < wave ? ws : 0;
            round += ws;
        }
        uint32_t e0 = before + incl - s;
        if (b <= n)
            u4[b / 4] = make_uint4(e0, e0 + c0, e0 + c0 + c1, e0 + c0 + c1 + c2);
        carry += round;
        cur = nxt;
    }
    if (tid == 0)
    {
        if (n % ROUND == 0)
            units[n] = carry;
        st->total = carry;
    }
}

/* LDS ordering among the lanes of one wave (its waves run independently) */
__device__ __forceinline__ void
wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* persistent waves over the pages: each wave stages its page in its own LDS
 * slot, lists the normal items and deforms them 64 at a time, while the
 * next page's loads are already in flight in registers */
template <int NC>
__global__ void __launch_bounds__(XHEAP_DEFORM_BLOCK)
xpu_heap_block_deform_kernel(xheap_params p)
{
    xheap_col regc[NC > 0 ? NC : 1];
    const xheap_col *cols = p.cols;
    if (NC > 0)
    {
#pragma unroll
        for (int j = 0; j < NC; j++)
            if ((uint32_t)j < p.natts_walk)
                regc[j] = p.cols[j];
        cols = regc;
    }
    constexpr uint32_t NW = XHEAP_DEFORM_BLOCK / 64;
    constexpr uint32_t NV = XPU_BLCKSZ / 8 / 64;
    __shared__ uint2 pg_all[NW][XPU_BLCKSZ / 8];
    __shared__ uint16_t items_all[NW][XPU_BLCKSZ / 4];
    const uint32_t wave = threadIdx.x / 64, lane = lane_id();
    uint2 *pg = pg_all[wave];
    uint16_t *items = items_all[wave];
    const uint8_t *pb = (const uint8_t *)pg;
    const uint32_t *lps = (const uint32_t *)(pb + XPU_PAGE_HEADER_SIZE);
    const uint32_t stride = gridDim.x * NW;
    bool bad = false;
    uint32_t page = blockIdx.x * NW + wave;
    uint2 v[NV];
    if (page < p.nunits)
    {
        const uint2 *gp = (const uint2 *)((const uint8_t *)p.src + p.block_offset +
                                          (uint64_t)page * XPU_BLCKSZ);
#pragma unroll
        for (uint32_t i = 0; i < NV; i++)
            v[i] = gp[lane + 64 * i];
    }
    for (; page < p.nunits; page += stride)
    {
        const uint64_t page_src = p.block_offset + (uint64_t)page * XPU_BLCKSZ;
        wave_sync();
#pragma unroll
        for (uint32_t i = 0; i < NV; i++)
            pg[lane + 64 * i] = v[i];
        if (page + stride < p.nunits)
        {
            const uint2 *gp = (const uint2 *)((const uint8_t *)p.src + page_src +
                                              (uint64_t)stride * XPU_BLCKSZ);
#pragma unroll
            for (uint32_t i = 0; i < NV; i++)
                v[i] = gp[lane + 64 * i];
        }
        wave_sync();
        uint32_t nlp = page_nlp((uint32_t)pb[XPU_PAGE_PD_LOWER] |
                                ((uint32_t)pb[XPU_PAGE_PD_LOWER + 1] << 8), &bad);
        uint32_t n = 0;
        for (uint32_t k0 = 0; k0 < nlp; k0 += 64)
        {
            uint32_t k = k0 + lane;
            bool ok = k < nlp && page_item_ok(lps[k], &bad);
            uint64_t m = __ballot(ok);
            if (ok)
                items[n + (uint32_t)__popcll(m & ((1ULL << lane) - 1))] = (uint16_t)k;
            n += (uint32_t)__popcll(m);
        }
        wave_sync();
        const uint32_t base = p.units[page];
        for (uint32_t b0 = 0; b0 < n; b0 += 64)
        {
            uint32_t idx = b0 + lane;
            bool active = idx < n && base + idx < p.capacity;
            uint32_t off = 0, len = 0;
            if (active)
            {
                uint32_t lp = lps[items[idx]];
                off = XPU_ITEMID_OFF(lp);
                len = XPU_ITEMID_LEN(lp);
            }
            deform_tuple<NC>(p, cols, pb, off, off + len, page_src, base + idx, base + b0, active,
                             &bad);
        }
    }
    if (__ballot(bad) && lane == 0)
        heap_error(p.st, XMSG_HEAP_TUPLE);
}

/* 64 row-index entries per 64-lane block: compact the visible rows onto the
 * first lanes, deform them from global memory */
template <int NC>
__global__ void __launch_bounds__(XHEAP_BLOCK)
xpu_heap_row_deform_kernel(xheap_params p)
{
    xheap_col regc[NC > 0 ? NC : 1];
    const xheap_col *cols = p.cols;
    if (NC > 0)
    {
#pragma unroll
        for (int j = 0; j < NC; j++)
            if ((uint32_t)j < p.natts_walk)
                regc[j] = p.cols[j];
        cols = regc;
    }
    __shared__ uint32_t slot[64];
    const uint32_t unit = blockIdx.x, lane = lane_id();
    bool bad = false;
    uint64_t hoff = 0;
    uint32_t hend = 0;
    uint32_t row_src = unit * 64 + lane;
    bool ok = row_src < p.nitems && row_item(p, row_src, &hoff, &hend, &bad);
    uint64_t m = __ballot(ok);
    if (ok)
        slot[__popcll(m & ((1ULL << lane) - 1))] = row_src;
    __syncthreads();
    const uint32_t n = (uint32_t)__popcll(m);
    const uint32_t base = p.units[unit];
    bool active = lane < n && base + lane < p.capacity;
    if (active)
        active = row_item(p, slot[lane], &hoff, &hend, &bad);
    const uint8_t *h = (const uint8_t *)p.src + hoff;
    deform_tuple<NC>(p, cols, h, 0, hend, hoff, base + lane, base, active, &bad);
    if (__ballot(bad) && lane == 0)
        heap_error(p.st, XMSG_HEAP_TUPLE);
}

extern "C" int
xpu_launch_heap_count(const xheap_params *hp, void *stream)
{
    if (hp->nunits > 0)
    {
        uint32_t per = XHEAP_COUNT_BLOCK / 64;
        hipLaunchKernelGGL(xpu_heap_count_kernel, dim3((hp->nunits + per - 1) / per),
                           dim3(XHEAP_COUNT_BLOCK), 0, (hipStream_t)stream, *hp);
    }
    hipLaunchKernelGGL(xpu_heap_scan_kernel, dim3(1), dim3(XHEAP_SCAN_BLOCK), 0,
                       (hipStream_t)stream, hp->units, hp->nunits, hp->st);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_heap_deform(const xheap_params *hp, int ncu, void *stream)
{
    if (hp->nunits == 0 || hp->capacity == 0)
        return 0;
    if (hp->format == KDS_FORMAT_BLOCK)
    {
        /* three 48 KB blocks of four waves per compute unit, or fewer when
         * the pages run out */
        uint32_t per = XHEAP_DEFORM_BLOCK / 64;
        uint32_t grid = (hp->nunits + per - 1) / per;
        if (grid > 3U * (uint32_t)ncu)
            grid = 3U * (uint32_t)ncu;
        if (hp->natts_walk <= XHEAP_REG_COLS)
            hipLaunchKernelGGL(xpu_heap_block_deform_kernel<XHEAP_REG_COLS>, dim3(grid),
                               dim3(XHEAP_DEFORM_BLOCK), 0, (hipStream_t)stream, *hp);
        else
            hipLaunchKernelGGL(xpu_heap_block_deform_kernel<0>, dim3(grid),
                               dim3(XHEAP_DEFORM_BLOCK), 0, (hipStream_t)stream, *hp);
    }
    else
    {
        if (hp->natts_walk <= XHEAP_REG_COLS)
            hipLaunchKernelGGL(xpu_heap_row_deform_kernel<XHEAP_REG_COLS>, dim3(hp->nunits),
                               dim3(XHEAP_BLOCK), 0, (hipStream_t)stream, *hp);
        else
            hipLaunchKernelGGL(xpu_heap_row_deform_kernel<0>, dim3(hp->nunits), dim3(XHEAP_BLOCK),
                               0, (hipStream_t)stream, *hp);
    }
    return (int)hipGetLastError();
}
