/*
 * xpu_sort.hip - GPU-Sort of a final buffer (the group-by results) by the
 * session's FuncOpCode__SortKeys descriptors, ahead of the write-back.
 *
 * The reference sorts the final buffer's row index with a bitonic network,
 * one launch per (scale, step) pair, each comparison re-reading both tuples
 * through kern_fetch_minimal_tuple_attr and the key type's comparator
 * (cuda_gpusort.cu:15-399, gpu_service.c:5199-5345).  Here every tuple is
 * read once:
 *   key pass      one lane per item walks its tuple once to the key attributes
 *                 and writes the key as 64-bit words whose unsigned order is
 *                 the comparator's order (NULLs first or last, ASC or DESC,
 *                 and an empty aggregate state sorting as NULL), plus the
 *                 least and greatest value of every word over all items
 *   radix passes  stable LSD radix sort of (words, item number) by 8-bit
 *                 digits of (word - least), last word first: subtracting a
 *                 word's least value keeps its order, and only the bytes of
 *                 (greatest - least) can differ, so a million int4 keys sort
 *                 in 3 passes, values of both signs near zero pay for their
 *                 range only, and a column without NULLs pays no pass for its
 *                 NULL class.  Each pass is a digit histogram per tile of
 *                 4,096 items, one exclusive scan over (digit, tile), and a
 *                 scatter that ranks equal digits with wave ballots
 *   permute       the sorted item numbers pick the row-index entries
 *
 * Key words (each key in turn, most significant first):
 *   32-bit and narrower values   one word: class << 32 | value bits
 *   64-bit values                class, then the 64-bit value bits
 *   numeric sums (int128)        class (with -inf / finite / +inf / NaN),
 *                                then the high and low halves of the sum
 *                                plus 2^63 (sums within int64 keep a
 *                                constant high half)
 * class is 1 for a value, 0 or 2 for NULL (NULLS FIRST / LAST).  Integers
 * have their sign bit flipped; floats have the sign bit set when positive
 * and every bit inverted when negative, -0 read as +0 (the reference's
 * comparators use < and >, which hold -0 and +0 equal) and NaN above every
 * value.  DESC inverts the value bits (and reverses the numeric classes),
 * never the NULL class.
 */
#include <hip/hip_runtime.h>
#include "pgstrom_xpu.h"
#include "xpu_kernels.h"

__device__ static inline uint64_t
sort_load64(const char *p)
{
    uint64_t v;
    __builtin_memcpy(&v, p, 8);
    return v;
}

__device__ static inline uint32_t
sort_enc_f16(uint32_t b)
{
    b &= 0xffffU;
    if ((b & 0x7fffU) > 0x7c00U)
        return 0xffffU;                             /* NaN */
    if ((b & 0x7fffU) == 0)
        b = 0;
    return (b & 0x8000U) ? (~b & 0xffffU) : (b | 0x8000U);
}

__device__ static inline uint32_t
sort_enc_f32(uint32_t b)
{
    if ((b & 0x7fffffffU) > 0x7f800000U)
        return 0xffffffffU;
    if ((b & 0x7fffffffU) == 0)
        b = 0;
    return (b & 0x80000000U) ? ~b : (b | 0x80000000U);
}

__device__ static inline uint64_t
sort_enc_f64(uint64_t b)
{
    if ((b & 0x7fffffffffffffffULL) > 0x7ff0000000000000ULL)
        return ~0ULL;
    if ((b & 0x7fffffffffffffffULL) == 0)
        b = 0;
    return (b & 0x8000000000000000ULL) ? ~b : (b | 0x8000000000000000ULL);
}

__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_sort_keys_kernel(xsort_params sp)
{
    __shared__ unsigned long long s_min[XSORT_MAX_WORDS];
    __shared__ unsigned long long s_max[XSORT_MAX_WORDS];
    const uint32_t nw = sp.nwords;
    if (threadIdx.x < nw)
    {
        s_min[threadIdx.x] = ~0ULL;
        s_max[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * XSORT_BLOCK + threadIdx.x;
    const bool valid = i < sp.n;
    /* one walk of the tuple finds every key attribute (the attribute walk
     * of kern_fetch_minimal_tuple_attr: NULL bitmap, alignment of fixed
     * values and of 4-byte varlena headers, 1- and 4-byte header lengths) */
    const char *kptr[XSORT_MAX_KEYS];
#pragma unroll
    for (int k = 0; k < XSORT_MAX_KEYS; k++)
        kptr[k] = nullptr;
    if (valid)
    {
        const kern_tupitem *t = (const kern_tupitem *)
            ((const char *)sp.kds + sp.kds->length - sp.rowindex[i]);
        const int natts = t->t_infomask2 & HEAP_NATTS_MASK;
        const int last = (int)sp.max_anum < natts ? (int)sp.max_anum : natts;
        const bool hasnull = (t->t_infomask & HEAP_HASNULL) != 0;
        const uint8_t *bits = (const uint8_t *)t + KERN_TUPITEM_T_BITS_OFFSET;
        const char *base = (const char *)t;
        uint32_t off = (uint32_t)t->t_hoff - MINIMAL_TUPLE_OFFSET;
        for (int j = 0; j < last; j++)
        {
            if (hasnull && !(bits[j >> 3] & (1 << (j & 7))))
                continue;                           /* NULL: kptr stays null */
            const kern_colmeta *cm = &sp.kds->colmeta[j];
            const uint32_t al = cm->attalign > 0 ? (uint32_t)cm->attalign : 1U;
            if (cm->attlen > 0 || base[off] == 0)
                off = (off + al - 1) & ~(al - 1);
#pragma unroll
            for (int k = 0; k < XSORT_MAX_KEYS; k++)
                if (k < (int)sp.nkeys && sp.keys[k].anum == j + 1)
                    kptr[k] = base + off;
            if (cm->attlen > 0)
                off += cm->attlen;
            else
            {
                const uint8_t b = (uint8_t)base[off];
                if (b & 0x01)
                    off += (b >> 1) & 0x7f;
                else
                {
                    uint32_t h;
                    __builtin_memcpy(&h, base + off, 4);
                    off += (h >> 2) & 0x3fffffffU;
                }
            }
        }
    }
    /* keys are wave-uniform: each key's words are stored and folded into
     * the least / greatest values as they are made */
#pragma unroll
    for (int k = 0; k < XSORT_MAX_KEYS; k++)
    {
        if (k >= (int)sp.nkeys)
            break;
        const xsort_key K = sp.keys[k];
        uint64_t kw[3] = {0, 0, 0};
        if (valid)
        {
            const char *p = kptr[k];
            const uint64_t null_class = K.nulls_first ? 0 : 2;
            uint64_t cls = 1, v = 0, v2 = 0;
            switch (K.kind)
            {
                case KSORT_KEY_KIND__VREF:
                    if (!p)
                        break;
                    switch (K.type_code)
                    {
                        case TypeOpCode__bool: v = *(const uint8_t *)p ? 1 : 0; break;
                        case TypeOpCode__int1: v = (uint32_t)(int32_t)*(const int8_t *)p ^ 0x80000000U; break;
                        case TypeOpCode__int2: v = (uint32_t)(int32_t)*(const int16_t *)p ^ 0x80000000U; break;
                        case TypeOpCode__int4:
                        case TypeOpCode__date: v = (uint32_t)*(const int32_t *)p ^ 0x80000000U; break;
                        case TypeOpCode__float2: v = sort_enc_f16(*(const uint16_t *)p); break;
                        case TypeOpCode__float4: v = sort_enc_f32(*(const uint32_t *)p); break;
                        case TypeOpCode__int8: v = sort_load64(p) ^ 0x8000000000000000ULL; break;
                        case TypeOpCode__float8: v = sort_enc_f64(sort_load64(p)); break;
                        default: break;
                    }
                    break;
                case KSORT_KEY_KIND__PMINMAX_INT64:
                    /* kagg_state__pminmax_int64_packed (cuda_gpusort.cu:66-93) */
                    if (p && (*(const uint32_t *)(p + 4) & __PAGG_MINMAX_ATTRS__VALID))
                        v = sort_load64(p + 8) ^ 0x8000000000000000ULL;
                    else
                        p = nullptr;
                    break;
                case KSORT_KEY_KIND__PMINMAX_FP64:
                    if (p && (*(const uint32_t *)(p + 4) & __PAGG_MINMAX_ATTRS__VALID))
                        v = sort_enc_f64(sort_load64(p + 8));
                    else
                        p = nullptr;
                    break;
                case KSORT_KEY_KIND__PSUM_INT64:
                    /* kagg_state__psum_int_packed: nitems > 0 (:124-151) */
                    if (p && (int64_t)sort_load64(p + 8) > 0)
                        v = sort_load64(p + 16) ^ 0x8000000000000000ULL;
                    else
                        p = nullptr;
                    break;
                case KSORT_KEY_KIND__PSUM_FP64:
                    if (p && (int64_t)sort_load64(p + 8) > 0)
                        v = sort_enc_f64(sort_load64(p + 16));
                    else
                        p = nullptr;
                    break;
                case KSORT_KEY_KIND__PAVG_INT64:
                case KSORT_KEY_KIND__PAVG_FP64:
                {
                    /* sum / nitems in float8 (__gpusort_prep_pavg_*, :401-459) */
                    const int64_t nitems = p ? (int64_t)sort_load64(p + 8) : 0;
                    if (nitems == 0)
                    {
                        p = nullptr;
                        break;
                    }
                    const uint64_t raw = sort_load64(p + 16);
                    double sum;
                    if (K.kind == KSORT_KEY_KIND__PAVG_INT64)
                        sum = (double)(int64_t)raw;
                    else
                        __builtin_memcpy(&sum, &raw, 8);
                    const double avg = sum / (double)nitems;
                    uint64_t bits;
                    __builtin_memcpy(&bits, &avg, 8);
                    v = sort_enc_f64(bits);
                    break;
                }
                case KSORT_KEY_KIND__PSUM_NUMERIC:
                {
                    /* kagg_state__psum_numeric_packed (:183-243): NaN sorts
                     * above +inf, -inf below every value */
                    if (!p || sort_load64(p + 8) == 0)
                    {
                        p = nullptr;
                        break;
                    }
                    const uint32_t attrs = *(const uint32_t *)(p + 4);
                    const uint32_t special = attrs & __PAGG_NUMERIC_ATTRS__MASK;
                    if (special == 0)
                    {
                        const uint64_t lo = sort_load64(p + 16);
                        const uint64_t lo_t = lo + 0x8000000000000000ULL;
                        const uint64_t hi_t = sort_load64(p + 24) + (lo_t < lo ? 1 : 0);
                        cls = 2;
                        v = hi_t ^ 0x8000000000000000ULL;       /* high half */
                        v2 = lo_t;                              /* low half */
                    }
                    else
                        cls = (special == __PAGG_NUMERIC_ATTRS__NINF) ? 1
                            : (special == __PAGG_NUMERIC_ATTRS__PINF) ? 3 : 4;
                    if (!K.order_asc)
                        cls = 5 - cls;
                    break;
                }
                default:
                    break;
            }
            const bool numeric = (K.kind == KSORT_KEY_KIND__PSUM_NUMERIC);
            if (!p)
            {
                cls = numeric ? (K.nulls_first ? 0 : 5) : null_class;
                v = v2 = 0;
            }
            else if (!K.order_asc)
            {
                v = ~v;
                v2 = ~v2;
            }
            if (K.nwords == 1)
                kw[0] = (cls << 32) | (v & 0xffffffffULL);
            else
            {
                kw[0] = cls;
                kw[1] = v;
                kw[2] = v2;
            }
            for (uint32_t q = 0; q < K.nwords; q++)
                sp.words[(uint64_t)(K.word + q) * sp.n + i] = kw[q];
        }
        for (uint32_t q = 0; q < K.nwords; q++)
        {
            const uint32_t wd = K.word + q;
            const uint64_t x = q == 0 ? kw[0] : q == 1 ? kw[1] : kw[2];
            unsigned long long lo = valid ? x : ~0ULL;
            unsigned long long hi = valid ? x : 0ULL;
            for (int d = 32; d >= 1; d >>= 1)
            {
                const unsigned long long a = __shfl_xor(lo, d);
                const unsigned long long b = __shfl_xor(hi, d);
                lo = a < lo ? a : lo;
                hi = b > hi ? b : hi;
            }
            if (__lane_id() == 0)
            {
                atomicMin(&s_min[wd], lo);
                atomicMax(&s_max[wd], hi);
            }
        }
    }
    if (valid)
        sp.idx[i] = i;
    __syncthreads();
    if (threadIdx.x < nw)
    {
        atomicMin((unsigned long long *)&sp.wmin[threadIdx.x], s_min[threadIdx.x]);
        atomicMax((unsigned long long *)&sp.wmax[threadIdx.x], s_max[threadIdx.x]);
    }
}

/* digit counts of one tile: counts[digit * ntiles + tile] */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_sort_hist_kernel(const uint64_t *words, uint32_t n, uint32_t word, uint64_t wbase,
                     uint32_t shift, uint32_t *counts)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t *key = words + (uint64_t)word * n;
    const uint32_t base = blockIdx.x * XSORT_TILE;
    for (int r = 0; r < XSORT_ROUNDS; r++)
    {
        const uint32_t j = base + r * XSORT_BLOCK + threadIdx.x;
        if (j < n)
            atomicAdd(&h[(uint32_t)((key[j] - wbase) >> shift) & 0xffU], 1U);
    }
    __syncthreads();
    counts[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

/* exclusive prefix of a 256-thread block's values (wave scans, then the
 * wave totals through LDS); *total gets the block's sum */
__device__ static inline uint32_t
sort_block_scan(uint32_t v, uint32_t *wsum, uint32_t *total)
{
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1)
    {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= (uint32_t)d)
            x += y;
    }
    if (lane == 63)
        wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < XSORT_BLOCK / 64; w++)
    {
        before += w < wave ? wsum[w] : 0;
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return x - v + before;
}

/* exclusive prefix over the tiles of one digit (block = digit): counts
 * [digit * ntiles + tile] become offsets within the digit, and
 * totals[digit] the digit's count */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_sort_scan_kernel(uint32_t *counts, uint32_t ntiles, uint32_t *totals)
{
    __shared__ uint32_t wsum[XSORT_BLOCK / 64];
    uint32_t *row = counts + (uint64_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += XSORT_BLOCK)
    {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < ntiles ? row[j] : 0;
        uint32_t total;
        const uint32_t ex = sort_block_scan(v, wsum, &total);
        if (j < ntiles)
            row[j] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0)
        totals[blockIdx.x] = carry;
}

/* stable scatter of one tile: items go out in (digit, tile, position) order.
 * A round takes 256 consecutive items, one per lane; lanes of a wave with
 * the same digit find each other with 8 ballots (one per digit bit), so an
 * item's place is the digit's running offset, the counts of the earlier
 * waves of the round, and its rank among its wave's equal digits. */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_sort_scatter_kernel(const uint64_t *words_in, const uint32_t *idx_in, uint64_t *words_out,
                        uint32_t *idx_out, uint32_t n, uint32_t nwords, uint32_t word,
                        uint64_t wbase, uint32_t shift, const uint32_t *counts,
                        const uint32_t *totals)
{
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[XSORT_BLOCK / 64][256];
    __shared__ uint32_t wsum[XSORT_BLOCK / 64];
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid / 64;
    const uint32_t lane = __lane_id();
    uint32_t all;
    const uint32_t digit_base = sort_block_scan(totals[tid], wsum, &all);
    run[tid] = digit_base + counts[(uint64_t)tid * gridDim.x + blockIdx.x];
    const uint64_t *key = words_in + (uint64_t)word * n;
    const uint32_t base = blockIdx.x * XSORT_TILE;
    for (int r = 0; r < XSORT_ROUNDS; r++)
    {
        for (uint32_t q = tid; q < (XSORT_BLOCK / 64) * 256; q += XSORT_BLOCK)
            wc[q / 256][q % 256] = 0;
        __syncthreads();
        const uint32_t j = base + r * XSORT_BLOCK + tid;
        const bool valid = j < n;
        const uint32_t d = valid ? (uint32_t)((key[j] - wbase) >> shift) & 0xffU : 0U;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++)
        {
            const bool bit = (d >> b) & 1U;
            const uint64_t bb = __ballot(valid && bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & ((1ULL << lane) - 1ULL));
        if (valid && rank == 0)
            wc[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid)
        {
            uint32_t pos = run[d] + rank;
            for (uint32_t w2 = 0; w2 < wave; w2++)
                pos += wc[w2][d];
            for (uint32_t k = 0; k < nwords; k++)
                words_out[(uint64_t)k * n + pos] = words_in[(uint64_t)k * n + j];
            idx_out[pos] = idx_in[j];
        }
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t w2 = 0; w2 < XSORT_BLOCK / 64; w2++)
            add += wc[w2][tid];
        run[tid] += add;
        __syncthreads();
    }
}

__global__ void
xpu_sort_permute_kernel(const uint64_t *rowindex, const uint32_t *idx, uint64_t *out, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = rowindex[idx[i]];
}

/* LIMIT, first step: the kept rows' tuple sizes (MAXALIGNed) in sorted
 * order, as exclusive offsets offs[j] and their total *used, so the caller
 * sizes the stage to exactly the kept bytes.  One block: a LIMIT keeps few
 * rows, and the scan carries a 64-bit running total across rounds. */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_sort_limit_size_kernel(const kern_data_store *kds, const uint64_t *ridx, uint32_t nout,
                           uint64_t *offs, uint64_t *used)
{
    __shared__ uint32_t wsum[XSORT_BLOCK / 64];
    const char *tail = (const char *)kds + kds->length;
    uint64_t carry = 0;
    for (uint32_t base = 0; base < nout; base += XSORT_BLOCK)
    {
        const uint32_t j = base + threadIdx.x;
        const kern_tupitem *t = j < nout ? (const kern_tupitem *)(tail - ridx[j]) : nullptr;
        const uint32_t sz = t ? ((t->t_len + 7U) & ~7U) : 0U;
        uint32_t total;
        const uint32_t ex = sort_block_scan(sz, wsum, &total);
        if (t)
            offs[j] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0)
        *used = carry;
}

/* LIMIT, second step: each kept tuple copied to the tail of the stage
 * (stage_len bytes, the total of the first step) in sorted order, with its
 * row id set to its sorted position, as kern_buffer_simple_limit packs them
 * (cuda_gpusort.cu:1401-1454, KERN_TUPITEM_SET_ROWID at :1446);
 * ridx_out[j] becomes the tuple's distance from the stage's end.  One lane
 * per kept row, any number of blocks. */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_sort_limit_copy_kernel(const kern_data_store *kds, const uint64_t *ridx, uint32_t nout,
                           const uint64_t *offs, char *stage, uint64_t stage_len,
                           uint64_t *ridx_out)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nout)
        return;
    const char *tail = (const char *)kds + kds->length;
    const kern_tupitem *t = (const kern_tupitem *)(tail - ridx[j]);
    const uint32_t t_len = t->t_len;
    const uint32_t sz = (t_len + 7U) & ~7U;
    const uint64_t end_off = offs[j] + sz;
    if (end_off > stage_len || t_len < offsetof(kern_tupitem, t_bits) + ROWID_SIZE)
        return;
    uint64_t *dst = (uint64_t *)(stage + stage_len - end_off);
    const uint64_t *src = (const uint64_t *)t;
    for (uint32_t q = 0; q < sz / 8; q++)
        dst[q] = src[q];
    const uint32_t rowid = j;
    __builtin_memcpy((char *)dst + t_len - ROWID_SIZE, &rowid, ROWID_SIZE);
    ridx_out[j] = end_off;
}

/* ------------------------------------------------------------------
 * Window rank over the sorted key words (kern_windowrank_exec_*,
 * cuda_gpusort.cu:814-987).  words is the sort's output, [nwords][n] in
 * sorted order; the partition keys are its first pw words and partition +
 * order keys its first ow words.  Rows of one partition are contiguous
 * (the partition keys lead the sort), so a row's partition start is the
 * first index on [0, i] whose partition words equal its own: a binary
 * search, as the reference's over its partition hashes.  Boundaries are
 * key changes; the reference's are changes of a 32-bit hash of the keys.
 * ------------------------------------------------------------------ */
__device__ static inline bool
wrank_equal(const uint64_t *words, uint32_t n, uint32_t nw, uint32_t a, uint32_t b)
{
    for (uint32_t w = 0; w < nw; w++)
        if (words[(uint64_t)w * n + a] != words[(uint64_t)w * n + b])
            return false;
    return true;
}

/* first index on [lo, i] equal to i on the first nw words */
__device__ static inline uint32_t
wrank_run_start(const uint64_t *words, uint32_t n, uint32_t nw, uint32_t lo, uint32_t i)
{
    uint32_t hi = i;
    while (lo < hi)
    {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (wrank_equal(words, n, nw, mid, i))
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

/* dense rank, first step: 1 where a peer group starts (a partition start
 * is one too: the order words include the partition words) */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_wrank_peer_kernel(const uint64_t *words, uint32_t n, uint32_t ow, uint32_t *flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        flag[i] = (i == 0 || !wrank_equal(words, n, ow, i - 1, i)) ? 1U : 0U;
}

/* keep[i] = 1 when row i's 0-based row_number / rank / dense_rank within
 * its partition is below lim (window_rank_limit - 1, unsigned as in the
 * reference); cpeer is the inclusive count of peer starts (dense rank) */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_wrank_keep_kernel(const uint64_t *words, uint32_t n, uint32_t pw, uint32_t ow, int func,
                      uint32_t lim, const uint32_t *cpeer, uint32_t *keep)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t part = wrank_run_start(words, n, pw, 0, i);
    uint32_t r;
    if (func == KSORT_WINDOW_FUNC__ROW_NUMBER)
        r = i - part;
    else if (func == KSORT_WINDOW_FUNC__RANK)
        r = wrank_run_start(words, n, ow, part, i) - part;
    else
        r = cpeer[i] - cpeer[part];
    keep[i] = r < lim ? 1U : 0U;
}

/* the kept entries of the sorted row index, in order: pos is the inclusive
 * prefix of keep */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_wrank_scatter_kernel(const uint64_t *ridx, const uint32_t *keep, const uint32_t *pos,
                         uint32_t n, uint64_t *ridx_out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && keep[i])
        ridx_out[pos[i] - 1] = ridx[i];
}

/* inclusive prefix of n counters, three steps: the sum of each tile of
 * XSORT_TILE, an exclusive scan of the tile sums in one block (its total
 * to *total), then each tile's scan from its base */
__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_scan_u32_tiles_kernel(const uint32_t *in, uint32_t n, uint32_t *tile_sum)
{
    __shared__ uint32_t wsum[XSORT_BLOCK / 64];
    const uint32_t base = blockIdx.x * XSORT_TILE;
    uint32_t v = 0;
    for (int r = 0; r < XSORT_ROUNDS; r++)
    {
        const uint32_t j = base + r * XSORT_BLOCK + threadIdx.x;
        v += j < n ? in[j] : 0U;
    }
    uint32_t total;
    (void)sort_block_scan(v, wsum, &total);
    if (threadIdx.x == 0)
        tile_sum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_scan_u32_top_kernel(uint32_t *tile_sum, uint32_t ntiles, uint32_t *total)
{
    __shared__ uint32_t wsum[XSORT_BLOCK / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += XSORT_BLOCK)
    {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < ntiles ? tile_sum[j] : 0U;
        uint32_t all;
        const uint32_t ex = sort_block_scan(v, wsum, &all);
        if (j < ntiles)
            tile_sum[j] = carry + ex;
        carry += all;
    }
    if (threadIdx.x == 0)
        *total = carry;
}

__global__ void __launch_bounds__(XSORT_BLOCK)
xpu_scan_u32_apply_kernel(const uint32_t *in, uint32_t n, const uint32_t *tile_base,
                          uint32_t *out)
{
    __shared__ uint32_t wsum[XSORT_BLOCK / 64];
    const uint32_t base = blockIdx.x * XSORT_TILE;
    uint32_t carry = tile_base[blockIdx.x];
    for (int r = 0; r < XSORT_ROUNDS; r++)
    {
        const uint32_t j = base + r * XSORT_BLOCK + threadIdx.x;
        const uint32_t v = j < n ? in[j] : 0U;
        uint32_t all;
        const uint32_t ex = sort_block_scan(v, wsum, &all);
        if (j < n)
            out[j] = carry + ex + v;
        carry += all;
    }
}

static int
scan_u32(const uint32_t *in, uint32_t n, uint32_t *tile_sum, uint32_t *total, uint32_t *out,
         hipStream_t st)
{
    const uint32_t ntiles = (n + XSORT_TILE - 1) / XSORT_TILE;
    hipLaunchKernelGGL(xpu_scan_u32_tiles_kernel, dim3(ntiles), dim3(XSORT_BLOCK), 0, st,
                       in, n, tile_sum);
    hipLaunchKernelGGL(xpu_scan_u32_top_kernel, dim3(1), dim3(XSORT_BLOCK), 0, st,
                       tile_sum, ntiles, total);
    hipLaunchKernelGGL(xpu_scan_u32_apply_kernel, dim3(ntiles), dim3(XSORT_BLOCK), 0, st,
                       in, n, (const uint32_t *)tile_sum, out);
    return (int)hipGetLastError();
}

/* Window rank of the n sorted rows: ridx_out gets the kept entries of ridx
 * (sorted order) and *nkept (device) their count.  Work arrays: flag, cpeer,
 * keep and pos of n counters each, tile_sum of ceil(n / XSORT_TILE). */
extern "C" int
xpu_launch_window_rank(const uint64_t *words, const uint64_t *ridx, uint32_t n, uint32_t pw,
                       uint32_t ow, int func, uint32_t limit, uint32_t *flag, uint32_t *cpeer,
                       uint32_t *keep, uint32_t *pos, uint32_t *tile_sum, uint32_t *nkept,
                       uint64_t *ridx_out, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (n == 0)
        return 0;
    const dim3 grid((n + XSORT_BLOCK - 1) / XSORT_BLOCK);
    int rc = 0;
    if (func == KSORT_WINDOW_FUNC__DENSE_RANK)
    {
        hipLaunchKernelGGL(xpu_wrank_peer_kernel, grid, dim3(XSORT_BLOCK), 0, st, words, n, ow, flag);
        rc = scan_u32(flag, n, tile_sum, nkept, cpeer, st);
    }
    if (rc == 0)
    {
        hipLaunchKernelGGL(xpu_wrank_keep_kernel, grid, dim3(XSORT_BLOCK), 0, st, words, n, pw, ow,
                           func, limit - 1U, (const uint32_t *)cpeer, keep);
        rc = scan_u32(keep, n, tile_sum, nkept, pos, st);
    }
    if (rc == 0)
        hipLaunchKernelGGL(xpu_wrank_scatter_kernel, grid, dim3(XSORT_BLOCK), 0, st, ridx,
                           (const uint32_t *)keep, (const uint32_t *)pos, n, ridx_out);
    return rc == 0 ? (int)hipGetLastError() : rc;
}

extern "C" int
xpu_launch_sort_limit_size(const kern_data_store *kds, const uint64_t *ridx, uint32_t nout,
                           uint64_t *offs, uint64_t *used, void *stream)
{
    hipLaunchKernelGGL(xpu_sort_limit_size_kernel, dim3(1), dim3(XSORT_BLOCK), 0,
                       (hipStream_t)stream, kds, ridx, nout, offs, used);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_sort_limit_copy(const kern_data_store *kds, const uint64_t *ridx, uint32_t nout,
                           const uint64_t *offs, char *stage, uint64_t stage_len,
                           uint64_t *ridx_out, void *stream)
{
    if (nout == 0)
        return 0;
    hipLaunchKernelGGL(xpu_sort_limit_copy_kernel, dim3((nout + XSORT_BLOCK - 1) / XSORT_BLOCK),
                       dim3(XSORT_BLOCK), 0, (hipStream_t)stream, kds, ridx, nout, offs, stage,
                       stage_len, ridx_out);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_sort_keys(const xsort_params *sp, void *stream)
{
    if (sp->n == 0)
        return 0;
    hipLaunchKernelGGL(xpu_sort_keys_kernel, dim3((sp->n + XSORT_BLOCK - 1) / XSORT_BLOCK),
                       dim3(XSORT_BLOCK), 0, (hipStream_t)stream, *sp);
    return (int)hipGetLastError();
}

/* one stable pass by the digit ((words[word] - base) >> shift) & 0xff;
 * counts holds 256 * (ceil(n / XSORT_TILE) + 1) entries */
extern "C" int
xpu_launch_sort_pass(const uint64_t *words_in, const uint32_t *idx_in, uint64_t *words_out,
                     uint32_t *idx_out, uint32_t n, uint32_t nwords, uint32_t word,
                     uint64_t base, uint32_t shift, uint32_t *counts, void *stream)
{
    if (n == 0)
        return 0;
    const uint32_t ntiles = (n + XSORT_TILE - 1) / XSORT_TILE;
    hipLaunchKernelGGL(xpu_sort_hist_kernel, dim3(ntiles), dim3(XSORT_BLOCK), 0,
                       (hipStream_t)stream, words_in, n, word, base, shift, counts);
    uint32_t *totals = counts + 256ULL * ntiles;
    hipLaunchKernelGGL(xpu_sort_scan_kernel, dim3(256), dim3(XSORT_BLOCK), 0, (hipStream_t)stream,
                       counts, ntiles, totals);
    hipLaunchKernelGGL(xpu_sort_scatter_kernel, dim3(ntiles), dim3(XSORT_BLOCK), 0,
                       (hipStream_t)stream, words_in, idx_in, words_out, idx_out, n, nwords,
                       word, base, shift, (const uint32_t *)counts, (const uint32_t *)totals);
    return (int)hipGetLastError();
}

extern "C" int
xpu_launch_sort_permute(const uint64_t *rowindex, const uint32_t *idx, uint64_t *out,
                        uint32_t n, void *stream)
{
    if (n == 0)
        return 0;
    hipLaunchKernelGGL(xpu_sort_permute_kernel, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, rowindex, idx, out, n);
    return (int)hipGetLastError();
}
