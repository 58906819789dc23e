/*
 * xpuhip_runtime.cpp - host side of the MI355X xPU library: the C ABI of
 * include/xpuhip.h over the HIP runtime.
 *
 * The roles it takes over from gpu_service.c:
 *   session open         gpuservHandleOpenSession (:3319-3410) + pointer
 *                        resolution (:3156-3317): here the blob is compiled
 *                        into an xplan and uploaded once;
 *   inner buffers        the kmrels import (:2197-2354): copied to HBM once
 *                        per session, dimension tables stay resident;
 *   task exec            __gpuservLaunchGpuTaskExecKernel (:3759-4003) with
 *                        its suspend/resume loop, here a resume position
 *                        returned to the caller;
 *   final write-back     gpuservHandleGpuTaskFinal (:5684-5882) and the
 *                        HASH->ROW conversion (:2780-2818).
 * Device memory comes from hipMalloc (coarse-grained HBM); control blocks are
 * cleared with a fill kernel, never hipMemsetAsync, so a caller may capture
 * the launch sequence in a hipGraph.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "xpuhip.h"
#include "xpu_compile.h"
#include "xpu_jit.h"
#include "xpu_kernels.h"

#define XERR(eb, code, msg) \
    xpu_set_error((eb), (code), "xpuhip_runtime.cpp", __LINE__, __func__, (msg))

/* Task scratch of xpuhip_task_exec_device_noalloc: the task control block,
 * then the projection's tile totals or the group-by's dump counters, dump
 * records and spill records.  Each device holds its own copy for the life of
 * the process.  Host code never passes this name to a kernel or a copy: it
 * is a host-side stand-in there; the device address comes from
 * hipGetSymbolAddress (g_scratch_ptr). */
#define XPU_NOALLOC_SCRATCH_BYTES   (32ULL << 20)
static __device__ __align__(256) char scratch[XPU_NOALLOC_SCRATCH_BYTES];

namespace {

struct Device
{
    bool        ready = false;
    int         id = 0;
    int         ncu = 0;
    size_t      lds_per_block = 0;
    std::string name;
};

Device      g_dev;
std::mutex  g_dev_lock;

/* device address of the no-alloc entry point's scratch array on each
 * device (set once per device with the device current) */
#define XPU_MAX_DEVICES         64
char       *g_scratch_ptr[XPU_MAX_DEVICES] = {};
thread_local double t_last_kernel_ms = 0.0;
thread_local xpuhip_task_stats t_last_stats;


bool hip_ok(hipError_t rc, kern_errorbuf *eb, const char *what);
bool scratch_symbol_address(char **p, kern_errorbuf *eb);

bool
hip_ok(hipError_t rc, kern_errorbuf *eb, const char *what)
{
    if (rc == hipSuccess)
        return true;
    char buf[200];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(rc));
    XERR(eb, ERRCODE_DEVICE_ERROR, buf);
    return false;
}

/* the device address of this device's scratch array; the device must be
 * current */
bool
scratch_symbol_address(char **p, kern_errorbuf *eb)
{
    void *ptr = nullptr;
    if (!hip_ok(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), eb, "hipGetSymbolAddress(scratch)"))
        return false;
    if (!ptr)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "hipGetSymbolAddress(scratch) returned no address");
        return false;
    }
    *p = (char *)ptr;
    return true;
}

int
ensure_device(kern_errorbuf *eb)
{
    std::lock_guard<std::mutex> g(g_dev_lock);
    if (g_dev.ready)
    {
        if (!hip_ok(hipSetDevice(g_dev.id), eb, "hipSetDevice"))
            return ERRCODE_DEVICE_ERROR;
        return 0;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no HIP device is available");
        return ERRCODE_DEVICE_ERROR;
    }
    hipDeviceProp_t prop;
    if (!hip_ok(hipSetDevice(g_dev.id), eb, "hipSetDevice") ||
        !hip_ok(hipGetDeviceProperties(&prop, g_dev.id), eb, "hipGetDeviceProperties"))
        return ERRCODE_DEVICE_ERROR;
    g_dev.ncu = prop.multiProcessorCount;
    g_dev.lds_per_block = prop.sharedMemPerBlock;
    g_dev.name = prop.gcnArchName;
    if (g_dev.id < 0 || g_dev.id >= XPU_MAX_DEVICES)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "device number out of range");
        return ERRCODE_DEVICE_ERROR;
    }
    if (!g_scratch_ptr[g_dev.id] && !scratch_symbol_address(&g_scratch_ptr[g_dev.id], eb))
        return ERRCODE_DEVICE_ERROR;
    g_dev.ready = true;
    return 0;
}

/* per-call device scratch: control block, look-back words, staging */
struct Scratch
{
    hipStream_t stream = nullptr;
    hipEvent_t  ev0 = nullptr, ev1 = nullptr, evm = nullptr;
    xpu_task_ctl *d_ctl = nullptr;
    xpu_task_ctl *h_ctl = nullptr;      /* pinned */
    uint64_t   *d_status = nullptr;
    size_t      status_cap = 0;
    uint64_t   *d_spill = nullptr;      /* group-by spill records */
    size_t      spill_cap = 0;          /* bytes */
    uint64_t   *d_dump = nullptr;       /* partial-state dump records */
    size_t      dump_cap = 0;           /* bytes */
    uint64_t   *d_dump_next = nullptr;  /* per-partition reservation counters */
    size_t      dump_next_cap = 0;
    char       *d_src = nullptr;
    size_t      src_cap = 0;
    char       *d_dst = nullptr;
    size_t      dst_cap = 0;
    uint32_t   *d_rec[2] = {};          /* multi-pass join records (odd, even levels) */
    size_t      rec_cap[2] = {};        /* bytes */
    uint32_t   *d_rec_cnt = nullptr;    /* [2][grid] records per region */
    size_t      rec_cnt_cap = 0;
    uint64_t   *d_block_status = nullptr;   /* scan + projection: block totals */
    size_t      block_status_cap = 0;
    /* heap-page / row chunks: their column tiles, the passes' status, column
     * table and unit counts, and host copies of what is uploaded */
    char       *d_tiles = nullptr;
    size_t      tiles_cap = 0;
    char       *d_hmeta = nullptr;
    size_t      hmeta_cap = 0;
    xheap_status *h_hstat = nullptr;    /* pinned */
    hipEvent_t  evh0 = nullptr, evh1 = nullptr;
    std::vector<char> h_tiles_head;
    std::vector<xheap_col> h_hcols;
};

bool
scratch_create(Scratch *s, kern_errorbuf *eb)
{
    return hip_ok(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), eb, "hipStreamCreate") &&
           hip_ok(hipEventCreate(&s->ev0), eb, "hipEventCreate") &&
           hip_ok(hipEventCreate(&s->ev1), eb, "hipEventCreate") &&
           hip_ok(hipEventCreate(&s->evm), eb, "hipEventCreate") &&
           hip_ok(hipMalloc((void **)&s->d_ctl, sizeof(xpu_task_ctl)), eb, "hipMalloc") &&
           hip_ok(hipHostMalloc((void **)&s->h_ctl, sizeof(xpu_task_ctl), hipHostMallocDefault),
                  eb, "hipHostMalloc") &&
           hip_ok(hipHostMalloc((void **)&s->h_hstat, sizeof(xheap_status), hipHostMallocDefault),
                  eb, "hipHostMalloc") &&
           hip_ok(hipEventCreate(&s->evh0), eb, "hipEventCreate") &&
           hip_ok(hipEventCreate(&s->evh1), eb, "hipEventCreate");
}

void
scratch_destroy(Scratch *s)
{
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->d_ctl) (void)hipFree(s->d_ctl);
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    if (s->d_status) (void)hipFree(s->d_status);
    if (s->d_spill) (void)hipFree(s->d_spill);
    if (s->d_dump) (void)hipFree(s->d_dump);
    if (s->d_dump_next) (void)hipFree(s->d_dump_next);
    if (s->d_src) (void)hipFree(s->d_src);
    if (s->d_dst) (void)hipFree(s->d_dst);
    for (int i = 0; i < 2; i++)
        if (s->d_rec[i]) (void)hipFree(s->d_rec[i]);
    if (s->d_rec_cnt) (void)hipFree(s->d_rec_cnt);
    if (s->d_block_status) (void)hipFree(s->d_block_status);
    if (s->d_tiles) (void)hipFree(s->d_tiles);
    if (s->d_hmeta) (void)hipFree(s->d_hmeta);
    if (s->h_hstat) (void)hipHostFree(s->h_hstat);
    if (s->evh0) (void)hipEventDestroy(s->evh0);
    if (s->evh1) (void)hipEventDestroy(s->evh1);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->evm) (void)hipEventDestroy(s->evm);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

bool
grow(char **p, size_t *cap, size_t need, kern_errorbuf *eb)
{
    if (*cap >= need)
        return true;
    if (*p)
        (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t sz = need + (need >> 3);
    if (!hip_ok(hipMalloc((void **)p, sz), eb, "hipMalloc"))
        return false;
    *cap = sz;
    return true;
}

}   /* namespace */

bool xpuhip_tuned_shape(const xplan *P, xpuhip_tuning *t);

struct xpuhip_session
{
    xpu_compiled    comp;
    xplan          *d_plan = nullptr;
    char           *d_const_pool = nullptr;
    kern_data_store *d_kds_dst_head = nullptr;
    kern_data_store *d_kds_final = nullptr;
    uint64_t        final_length = 0;
    /* group-by overflow table: the groups a launch finds no room for in
     * kds_final, merged in after the launch once kds_final has grown */
    kern_data_store *d_kds_ovf = nullptr;
    uint64_t        ovf_length = 0;
    /* CPU-fallback rows of the session's tasks (KDS_FORMAT_FALLBACK), until
     * xpuhip_session_fallback takes them */
    kern_data_store *d_kds_fallback = nullptr;
    /* launches on kds_final hold it shared; growing kds_final holds it
     * exclusively (the reference's m_kds_rwlock, gpu_service.c:2423) */
    std::shared_mutex final_rw;     /* also held around the fallback buffer */
    char           *d_kmrels = nullptr;
    uint64_t       *d_bloom = nullptr;      /* Bloom filters of the inner relations */
    uint64_t        bloom_off[XP_MAX_DEPTH + 1] = {};   /* words from d_bloom */
    uint32_t        bloom_mask[XP_MAX_DEPTH + 1] = {};  /* words - 1 (0: no filter) */
    uint32_t        maxdup[XP_MAX_DEPTH + 1] = {};      /* most items sharing a join hash */
    uint64_t        bloom_lds_off = 0;      /* the depth-1 filter folded for the screen's LDS */
    uint32_t        bloom_lds_words = 0;
    int64_t        *d_params = nullptr;     /* XO_PARAM values */
    uint32_t       *d_xids = nullptr;       /* parallelCurrentXids (row-chunk visibility) */
    struct PinnedBuf *pinned = nullptr;     /* pinned hash results: the shared buffer */
    char           *d_inner_own[XP_MAX_DEPTH + 1] = {};  /* inner buffers rebuilt from pinned
                                                            results (owned) */
    bool            inner_ready = false;
    char           *d_sort = nullptr;       /* GPU-Sort work space of task_final */
    size_t          sort_cap = 0;
    char           *d_sort_stage = nullptr; /* GPU-Sort LIMIT: the kept tuples, packed */
    size_t          sort_stage_cap = 0;
    xpuhip_tuning   tuning;
    std::mutex      lock;
    std::vector<Scratch *> pool;
    uint64_t        ntasks = 0;
};

/* Pinned inner buffers (gpu_service.c:2099-2196): the KDS_FORMAT_HASH
 * buffer that the projections of every session of one query_plan_id append
 * to, kept while a session of that plan is open.  A join names it by the
 * plan id in its kern_multirels chunk (buffer_id) and attaches a copy. */
struct PinnedBuf
{
    uint64_t    plan_id = 0;
    char       *d_kds = nullptr;
    uint64_t    length = 0;
    int         refs = 0;
};

namespace {

std::mutex g_pinned_lock;
std::map<uint64_t, PinnedBuf *> g_pinned;

/* Task scratch (stream, events, pinned control block, grown buffers) of
 * closed sessions, kept for the next sessions: opening and closing a
 * session per query then costs no stream or pinned-memory setup. */
std::mutex g_idle_lock;
std::vector<Scratch *> g_idle;
#define XPU_IDLE_SCRATCH        64
#define XPU_IDLE_KEEP_BYTES     (256ULL << 20)

void
scratch_trim(Scratch *x)
{
    auto trim = [](char **p, size_t *cap) {
        if (*p && *cap > XPU_IDLE_KEEP_BYTES)
        {
            (void)hipFree(*p);
            *p = nullptr;
            *cap = 0;
        }
    };
    trim((char **)&x->d_status, &x->status_cap);
    trim((char **)&x->d_spill, &x->spill_cap);
    trim((char **)&x->d_dump, &x->dump_cap);
    trim(&x->d_src, &x->src_cap);
    trim(&x->d_dst, &x->dst_cap);
    trim(&x->d_tiles, &x->tiles_cap);
    for (int i = 0; i < 2; i++)
        trim((char **)&x->d_rec[i], &x->rec_cap[i]);
}

/* a closed session's scratch, to the idle list (or freed beyond it) */
void
scratch_retire(Scratch *x)
{
    (void)hipStreamSynchronize(x->stream);
    scratch_trim(x);
    {
        std::lock_guard<std::mutex> g(g_idle_lock);
        if (g_idle.size() < XPU_IDLE_SCRATCH)
        {
            g_idle.push_back(x);
            return;
        }
    }
    scratch_destroy(x);
}

Scratch *
scratch_get(xpuhip_session *s, kern_errorbuf *eb)
{
    {
        std::lock_guard<std::mutex> g(s->lock);
        if (!s->pool.empty())
        {
            Scratch *x = s->pool.back();
            s->pool.pop_back();
            return x;
        }
    }
    {
        std::lock_guard<std::mutex> g(g_idle_lock);
        if (!g_idle.empty())
        {
            Scratch *x = g_idle.back();
            g_idle.pop_back();
            return x;
        }
    }
    Scratch *x = new Scratch();
    if (!scratch_create(x, eb))
    {
        scratch_destroy(x);
        return nullptr;
    }
    return x;
}

void
scratch_put(xpuhip_session *s, Scratch *x)
{
    std::lock_guard<std::mutex> g(s->lock);
    s->pool.push_back(x);
}

size_t
lds_bytes_for(const xplan *P, uint32_t nents)
{
    return (size_t)nents * P->lds_entry_bytes;
}

/* plans without joins run the scan + group-by kernel */
bool
uses_scan_agg(const xpuhip_session *s)
{
    const xplan *P = &s->comp.plan;
    return s->comp.has_groupby && P->nrels == 0 && P->fused;
}

/* largest LDS table (entries) that fits in budget bytes */
uint32_t
lds_nents_within(const xplan *P, size_t budget)
{
    if (!P->lds_fast || P->lds_entry_bytes == 0)
        return 0;
    size_t n = budget / P->lds_entry_bytes;
    return (uint32_t)(n > (1U << 20) ? (1U << 20) : n);
}

/* default launch shape per task kind (the autotuner overrides it) */
void
default_tuning(const xpuhip_session *s, xpuhip_tuning *t)
{
    const xplan *P = &s->comp.plan;
    memset(t, 0, sizeof(*t));
    if (uses_scan_agg(s))
    {
        /* one 1024-thread block per CU owning (nearly) all of the CU's LDS as
         * its partial table; 4 rows per lane per wave tile (C1, 100M rows:
         * scan pass 1.23 ms against 1.55 ms with 512 threads; the
         * interpreter's kernels run 512) */
        t->block_size = XP_AGG_BLOCK_MAX;
        t->rows_per_thread = 4;
        t->blocks_per_cu = 1;
        size_t budget = g_dev.lds_per_block > 4096 ? g_dev.lds_per_block - 4096 : 0;
        t->lds_nents = lds_nents_within(P, budget);
    }
    else if (s->comp.has_groupby)
    {
        /* joins + group-by: two 512-thread blocks per CU (measured on the
         * 200M-row star join: 15 ms against 25 ms with 256-thread blocks);
         * a wave screens 8 row batches of a tile before its joins run over
         * the rows that pass */
        t->block_size = 512;
        t->rows_per_thread = 8;
        t->blocks_per_cu = 2;
        t->lds_nents = lds_nents_within(P, 48 * 1024);
    }
    else
    {
        /* projections (count, tile scan, write passes over wave tiles):
         * 256-thread blocks; blocks per CU by chunk size (0) */
        t->block_size = 256;
        t->rows_per_thread = 1;
        t->blocks_per_cu = 0;
    }
}

/* host image of the single group tuple of a no-GROUP-BY kds_final
 * (states in their initial values, cuda_gpupreagg.cu:18-223) */
std::vector<char>
nogroup_tuple(const xplan *P, uint32_t *p_unit)
{
    uint32_t t_hoff = P->gb_t_hoff;
    uint32_t t_len = t_hoff + P->gb_states_end + ROWID_SIZE;
    uint32_t unit = XPU_TYPEALIGN(XPU_L1_CACHELINE_SZ, offsetof(kern_hashitem, t) + t_len);
    std::vector<char> buf(unit, 0);
    kern_hashitem *hi = (kern_hashitem *)buf.data();
    char *tb = (char *)&hi->t;
    uint16_t infomask = HEAP_HASNULL;
    for (uint32_t j = 0; j < P->gb_nattrs; j++)
    {
        const xagg *ag = &P->aggs[j];
        char *s = tb + t_hoff + ag->state_off;
        hi->t.t_bits[j >> 3] |= (uint8_t)(1 << (j & 7));
        int64_t v;
        double d;
        switch (ag->action)
        {
            case KAGG_ACTION__NROWS_ANY:
            case KAGG_ACTION__NROWS_COND:
                break;
            case KAGG_ACTION__PMIN_INT32: case KAGG_ACTION__PMIN_INT64:
                *(int32_t *)s = ag->state_size << 2;
                v = INT64_MAX; memcpy(s + 8, &v, 8);
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PMAX_INT32: case KAGG_ACTION__PMAX_INT64:
                *(int32_t *)s = ag->state_size << 2;
                v = INT64_MIN; memcpy(s + 8, &v, 8);
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PMIN_FP64:
                *(int32_t *)s = ag->state_size << 2;
                d = 1.7976931348623157e308; memcpy(s + 8, &d, 8);
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PMAX_FP64:
                *(int32_t *)s = ag->state_size << 2;
                d = -1.7976931348623157e308; memcpy(s + 8, &d, 8);
                infomask |= HEAP_HASVARWIDTH;
                break;
            case KAGG_ACTION__PSUM_NUMERIC: case KAGG_ACTION__PAVG_NUMERIC:
            {
                *(int32_t *)s = ag->state_size << 2;
                const uint32_t attrs = ag->num_scale;
                memcpy(s + 4, &attrs, 4);
                infomask |= HEAP_HASVARWIDTH;
                break;
            }
            default:
                *(int32_t *)s = ag->state_size << 2;
                infomask |= HEAP_HASVARWIDTH;
                break;
        }
    }
    hi->t.t_len = t_len;
    hi->t.t_infomask2 = (uint16_t)(P->gb_nattrs & HEAP_NATTS_MASK);
    hi->t.t_infomask = infomask;
    hi->t.t_hoff = (uint8_t)(t_hoff + MINIMAL_TUPLE_OFFSET);
    *p_unit = unit;
    return buf;
}

/* (re)initialises the session's kds_final on the device */
int
init_kds_final(xpuhip_session *s, kern_errorbuf *eb)
{
    const kern_data_store *head = (const kern_data_store *)s->comp.kds_final_head.data();
    uint64_t head_len = s->comp.kds_final_head.size();
    const xplan *P = &s->comp.plan;
    std::vector<char> img(head_len);
    memcpy(img.data(), head, head_len);
    kern_data_store *h = (kern_data_store *)img.data();
    uint64_t length = s->final_length;

    h->length = length;
    h->usage = 0;
    h->nitems = 0;
    if (P->nkeys == 0 && !s->comp.pinned_hash)
    {
        /* ROW-format kds_final holding the single group */
        uint32_t unit;
        std::vector<char> tup = nogroup_tuple(P, &unit);
        h->format = KDS_FORMAT_ROW;
        h->hash_nslots = 0;
        h->nitems = 1;
        h->usage = unit;
        uint64_t rowidx = unit - offsetof(kern_hashitem, t);
        if (!hip_ok(hipMemcpy(s->d_kds_final, img.data(), head_len, hipMemcpyHostToDevice), eb, "hipMemcpy") ||
            !hip_ok(hipMemcpy((char *)s->d_kds_final + head_len, &rowidx, 8, hipMemcpyHostToDevice), eb, "hipMemcpy") ||
            !hip_ok(hipMemcpy((char *)s->d_kds_final + length - unit, tup.data(), unit,
                              hipMemcpyHostToDevice), eb, "hipMemcpy"))
            return ERRCODE_DEVICE_ERROR;
        return 0;
    }
    if (!hip_ok(hipMemcpy(s->d_kds_final, img.data(), head_len, hipMemcpyHostToDevice), eb, "hipMemcpy"))
        return ERRCODE_DEVICE_ERROR;
    xpu_launch_fill((char *)s->d_kds_final + head_len, 0, h->hash_nslots, nullptr);
    if (!hip_ok(hipDeviceSynchronize(), eb, "kds_final init"))
        return ERRCODE_DEVICE_ERROR;
    return 0;
}

const char *
message_of(const xpuhip_session *s, uint32_t msgid)
{
    switch (msgid)
    {
        case XMSG_LOOKBACK_TIMEOUT: return "tile prefix look-back did not complete";
        case XMSG_FINAL_NO_SPACE: return "GpuPreAgg(GroupBy) - no space to write on kds_final";
        case XMSG_TEXT_EXTERNAL: return "text datum is compressed or external";
        case XMSG_GROUP_RETRY: return "group-by slot lock retry limit exceeded";
        case XMSG_TILE_TOO_LARGE: return "one tile of rows does not fit into kds_dst";
        case XMSG_JOIN_RECORDS: return "join records exceed their buffer";
        case XMSG_HEAP_ITEM: return "heap page line pointer or row-index entry out of its chunk";
        case XMSG_HEAP_TUPLE: return "heap tuple attribute beyond its item";
        case XMSG_ARROW_TYPE: return "not a mappable Arrow data type";
        case XMSG_FALLBACK_FULL: return "CPU fallback buffer is full";
        case XMSG_NUM_RANGE: return "numeric value is out of range";
        case XMSG_NUM_HEADER: return "corrupted numeric header";
        case XMSG_NUM_SPECIAL: return "unknown special numeric value";
        case XMSG_NUM_DIV0: return "division by zero";
        case XMSG_NUM_INT_RANGE: return "integer out of range";
        case XMSG_NUM_NANINF: return "cannot convert NaN/Inf to integer";
        case XMSG_NUM_FLOAT_RANGE: return "float out of range";
    }
    if (msgid < s->comp.messages.size())
        return s->comp.messages[msgid].c_str();
    return "device error";
}

}   /* namespace */

/* ====================================================================
 * C ABI
 * ==================================================================== */
extern "C" int
xpuhip_init(int device, kern_errorbuf *eb)
{
    {
        std::lock_guard<std::mutex> g(g_dev_lock);
        if (!g_dev.ready)
            g_dev.id = device < 0 ? 0 : device;
    }
    return ensure_device(eb);
}

extern "C" void
xpuhip_shutdown(void)
{
    std::lock_guard<std::mutex> g(g_dev_lock);
    if (g_dev.ready)
        (void)hipDeviceSynchronize();
    g_dev.ready = false;
}

extern "C" int
xpuhip_session_check(const kern_session_info *session, size_t length, kern_errorbuf *eb)
{
    xpu_compiled *c = new xpu_compiled();
    int rc = xpu_compile_session(session, length, c, eb);
    delete c;
    return rc;
}

extern "C" int
xpuhip_jit_prepare(const kern_session_info *session, size_t length, kern_errorbuf *eb)
{
    std::unique_ptr<xpu_compiled> c(new xpu_compiled());
    int rc = xpu_compile_session(session, length, c.get(), eb);
    if (rc)
        return rc;
    const xplan *P = &c->plan;
    const bool scan_agg = c->has_groupby && P->nrels == 0 && P->fused;
    const int kind = scan_agg ? XJ_SCAN_AGG : XJ_PIPELINE;
    const int param = scan_agg ? 4 : (c->has_groupby ? 1 : 0);
    std::string log;
    if (!xpu_jit_code(P, kind, param, &log))
    {
        if (log.size() > 150)
            log.resize(150);
        XERR(eb, ERRCODE_DEVICE_ERROR, ("specialized kernel does not compile: " + log).c_str());
        return ERRCODE_DEVICE_ERROR;
    }
    return 0;
}

extern "C" int64_t
xpuhip_session_explain(const kern_session_info *session, size_t length, char *buf, size_t buflen,
                       kern_errorbuf *eb)
{
    xpu_compiled *c = new xpu_compiled();
    int rc = xpu_compile_session(session, length, c, eb);
    if (rc)
    {
        delete c;
        return -1;
    }
    std::string s = xpu_plan_dump(c);
    delete c;
    if (buf && buflen > 0)
    {
        size_t n = s.size() < buflen - 1 ? s.size() : buflen - 1;
        memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return (int64_t)s.size();
}

extern "C" int
xpuhip_session_open(const kern_session_info *session, size_t length,
                    xpuhip_session **p_sess, kern_errorbuf *eb)
{
    int rc;
    *p_sess = nullptr;
    if ((rc = ensure_device(eb)) != 0)
        return rc;
    xpuhip_session *s = new xpuhip_session();
    if ((rc = xpu_compile_session(session, length, &s->comp, eb)) != 0)
    {
        delete s;
        return rc;
    }
    default_tuning(s, &s->tuning);
    {
        /* a shape tuned earlier in this process */
        xpuhip_tuning tt;
        if (xpuhip_tuned_shape(&s->comp.plan, &tt))
            s->tuning = tt;
    }
    /* plan and constants */
    size_t pool_len = s->comp.const_pool.size() ? s->comp.const_pool.size() : 8;
    if (!hip_ok(hipMalloc((void **)&s->d_const_pool, pool_len), eb, "hipMalloc") ||
        !hip_ok(hipMalloc((void **)&s->d_plan, sizeof(xplan)), eb, "hipMalloc"))
    {
        xpuhip_session_close(s);
        return ERRCODE_DEVICE_ERROR;
    }
    if (!s->comp.const_pool.empty() &&
        !hip_ok(hipMemcpy(s->d_const_pool, s->comp.const_pool.data(), s->comp.const_pool.size(),
                          hipMemcpyHostToDevice), eb, "hipMemcpy"))
    {
        xpuhip_session_close(s);
        return ERRCODE_DEVICE_ERROR;
    }
    s->comp.plan.const_pool = (uint64_t)(uintptr_t)s->d_const_pool;
    {
        size_t np = s->comp.params.size() ? s->comp.params.size() : 1;
        if (!hip_ok(hipMalloc((void **)&s->d_params, 8 * np), eb, "hipMalloc") ||
            (!s->comp.params.empty() &&
             !hip_ok(hipMemcpy(s->d_params, s->comp.params.data(), 8 * s->comp.params.size(),
                               hipMemcpyHostToDevice), eb, "hipMemcpy")))
        {
            xpuhip_session_close(s);
            return ERRCODE_DEVICE_ERROR;
        }
    }
    if (!s->comp.xids.empty() &&
        (!hip_ok(hipMalloc((void **)&s->d_xids, 4 * s->comp.xids.size()), eb, "hipMalloc") ||
         !hip_ok(hipMemcpy(s->d_xids, s->comp.xids.data(), 4 * s->comp.xids.size(),
                           hipMemcpyHostToDevice), eb, "hipMemcpy")))
    {
        xpuhip_session_close(s);
        return ERRCODE_DEVICE_ERROR;
    }
    if (!s->comp.kds_dst_head.empty())
    {
        if (!hip_ok(hipMalloc((void **)&s->d_kds_dst_head, s->comp.kds_dst_head.size()), eb, "hipMalloc") ||
            !hip_ok(hipMemcpy(s->d_kds_dst_head, s->comp.kds_dst_head.data(), s->comp.kds_dst_head.size(),
                              hipMemcpyHostToDevice), eb, "hipMemcpy"))
        {
            xpuhip_session_close(s);
            return ERRCODE_DEVICE_ERROR;
        }
    }
    if (s->comp.has_groupby)
    {
        const kern_data_store *fh = (const kern_data_store *)s->comp.kds_final_head.data();
        uint64_t len = fh->length;
        uint64_t need = s->comp.kds_final_head.size() + 8ULL * fh->hash_nslots + (1ULL << 20);
        if (len < need)
            len = need + (64ULL << 20);
        len &= ~(uint64_t)(XPU_L1_CACHELINE_SZ - 1);
        s->final_length = len;
        if (!hip_ok(hipMalloc((void **)&s->d_kds_final, len), eb, "hipMalloc(kds_final)") ||
            (rc = init_kds_final(s, eb)) != 0)
        {
            xpuhip_session_close(s);
            return rc ? rc : ERRCODE_DEVICE_ERROR;
        }
    }
    if (!s->comp.kds_fallback_head.empty())
    {
        /* the CPU-fallback buffer, as long as the session's header asks
         * (__allocateGpuFallbackBuffer, gpu_service.c:2556-2585) */
        std::vector<char> img = s->comp.kds_fallback_head;
        kern_data_store *fh = (kern_data_store *)img.data();
        fh->nitems = 0;
        fh->usage = 0;
        fh->length &= ~(uint64_t)(XPU_MAXIMUM_ALIGNOF - 1);
        if (!hip_ok(hipMalloc((void **)&s->d_kds_fallback, fh->length), eb, "hipMalloc(kds_fallback)") ||
            !hip_ok(hipMemcpy(s->d_kds_fallback, img.data(), img.size(), hipMemcpyHostToDevice), eb,
                    "hipMemcpy"))
        {
            xpuhip_session_close(s);
            return ERRCODE_DEVICE_ERROR;
        }
    }
    if (s->comp.pinned_hash)
    {
        /* the plan's pinned buffer: shared by its sessions, created by the
         * first one */
        std::lock_guard<std::mutex> g(g_pinned_lock);
        const uint64_t plan_id = session->query_plan_id;
        auto it = g_pinned.find(plan_id);
        PinnedBuf *pb = it != g_pinned.end() ? it->second : nullptr;
        if (!pb)
        {
            /* the header's length, as the reference allocates the pieces of a
             * pinned buffer (executor.c:664-668), at least its slots */
            const kern_data_store *fh = (const kern_data_store *)s->comp.kds_final_head.data();
            uint64_t len = fh->length;
            uint64_t need = s->comp.kds_final_head.size() + 8ULL * fh->hash_nslots + (64ULL << 10);
            if (len < need)
                len = need;
            len = (len + XPU_L1_CACHELINE_SZ - 1) & ~(uint64_t)(XPU_L1_CACHELINE_SZ - 1);
            s->final_length = len;
            if (!hip_ok(hipMalloc((void **)&s->d_kds_final, len), eb, "hipMalloc(pinned buffer)") ||
                (rc = init_kds_final(s, eb)) != 0)
            {
                xpuhip_session_close(s);
                return rc ? rc : ERRCODE_DEVICE_ERROR;
            }
            pb = new PinnedBuf();
            pb->plan_id = plan_id;
            pb->d_kds = (char *)s->d_kds_final;
            pb->length = len;
            g_pinned[plan_id] = pb;
        }
        pb->refs++;
        s->pinned = pb;
        s->d_kds_final = (kern_data_store *)pb->d_kds;
        s->final_length = pb->length;
    }
    if (!hip_ok(hipMemcpy(s->d_plan, &s->comp.plan, sizeof(xplan), hipMemcpyHostToDevice), eb, "hipMemcpy"))
    {
        xpuhip_session_close(s);
        return ERRCODE_DEVICE_ERROR;
    }
    s->inner_ready = (s->comp.plan.nrels == 0);
    *p_sess = s;
    return 0;
}

/* A copy of the pinned buffer of plan_id with hash slots for its rows
 * (KDS_GET_HASHSLOT_WIDTH), owned by the attaching session: the row index
 * and the tuple area are copied, the items re-chained
 * (gpu_service.c:1906-2056 with kern_buffer_reconstruction). */
static int
attach_pinned(xpuhip_session *s, uint32_t d, uint64_t plan_id, uint64_t *nitems_out,
              uint64_t *nslots_out, kern_errorbuf *eb)
{
    PinnedBuf *pb = nullptr;
    {
        std::lock_guard<std::mutex> g(g_pinned_lock);
        auto it = g_pinned.find(plan_id);
        if (it != g_pinned.end())
            pb = it->second;
    }
    char msg[128];
    if (!pb)
    {
        snprintf(msg, sizeof(msg), "pinned inner buffer %llx was not found",
                 (unsigned long long)plan_id);
        XERR(eb, ERRCODE_DEVICE_ERROR, msg);
        return ERRCODE_DEVICE_ERROR;
    }
    if (!hip_ok(hipDeviceSynchronize(), eb, "pinned buffer"))
        return ERRCODE_DEVICE_ERROR;
    kern_data_store h0;
    if (!hip_ok(hipMemcpy(&h0, pb->d_kds, sizeof(h0), hipMemcpyDeviceToHost), eb, "hipMemcpy"))
        return ERRCODE_DEVICE_ERROR;
    std::vector<char> head(KDS_HEAD_LENGTH(&h0));
    if (!hip_ok(hipMemcpy(head.data(), pb->d_kds, head.size(), hipMemcpyDeviceToHost), eb, "hipMemcpy"))
        return ERRCODE_DEVICE_ERROR;
    kern_data_store *h = (kern_data_store *)head.data();
    const uint64_t hl = head.size();
    const uint64_t n = h->nitems, usage = h->usage, src_slots = h->hash_nslots;
    const uint64_t nslots = KDS_GET_HASHSLOT_WIDTH(n);
    const uint64_t length = XPU_MAXALIGN(hl + 8 * nslots + 8 * n + usage);
    if (length >= (1ULL << 32))
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "pinned inner buffer is 4GB or larger");
        return ERRCODE_DEVICE_ERROR;
    }
    h->length = length;
    h->hash_nslots = nslots;
    if (s->d_inner_own[d])
        (void)hipFree(s->d_inner_own[d]);
    s->d_inner_own[d] = nullptr;
    char *dst = nullptr;
    if (!hip_ok(hipMalloc((void **)&dst, length), eb, "hipMalloc(inner buffer)"))
        return ERRCODE_DEVICE_ERROR;
    s->d_inner_own[d] = dst;
    if (!hip_ok(hipMemcpy(dst, head.data(), hl, hipMemcpyHostToDevice), eb, "hipMemcpy") ||
        !hip_ok((hipError_t)xpu_launch_fill(dst + hl, 0, nslots, nullptr), eb, "fill") ||
        (n > 0 && !hip_ok(hipMemcpy(dst + hl + 8 * nslots, pb->d_kds + hl + 8 * src_slots, 8 * n,
                                    hipMemcpyDeviceToDevice), eb, "hipMemcpy")) ||
        (usage > 0 && !hip_ok(hipMemcpy(dst + length - usage, pb->d_kds + pb->length - usage, usage,
                                        hipMemcpyDeviceToDevice), eb, "hipMemcpy")) ||
        !hip_ok((hipError_t)xpu_launch_hash_rechain((kern_data_store *)dst, (uint32_t)n, nullptr),
                eb, "rechain") ||
        !hip_ok(hipDeviceSynchronize(), eb, "pinned buffer copy"))
        return ERRCODE_DEVICE_ERROR;
    *nitems_out = n;
    *nslots_out = nslots;
    return 0;
}

extern "C" int
xpuhip_session_attach_inner(xpuhip_session *s, const kern_multirels *kmrels, size_t length,
                            kern_errorbuf *eb)
{
    xplan *P = &s->comp.plan;
    if (!kmrels || length < offsetof(kern_multirels, chunks) ||
        kmrels->num_inner_rels < P->nrels ||
        length < offsetof(kern_multirels, chunks) + sizeof(kern_multirels_chunk) * kmrels->num_inner_rels)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "kern_multirels does not match the session's joins");
        return ERRCODE_DEVICE_ERROR;
    }
    int rc;
    if ((rc = ensure_device(eb)) != 0)
        return rc;
    for (uint32_t d = 1; d <= P->nrels; d++)
    {
        const kern_multirels_chunk *ch = &kmrels->chunks[d - 1];
        if (ch->is_nestloop || ch->right_outer || ch->gist_offset)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "nest-loop, GiST and RIGHT OUTER joins are not supported");
            return ERRCODE_DEVICE_ERROR;
        }
        if (ch->pinned_buffer)
            continue;       /* a pinned inner buffer on the device, by buffer_id */
        if (ch->kds_offset == 0 || ch->kds_offset + sizeof(kern_data_store) > length)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "inner KDS out of kern_multirels");
            return ERRCODE_DEVICE_ERROR;
        }
        const kern_data_store *k = (const kern_data_store *)((const char *)kmrels + ch->kds_offset);
        if (k->format != KDS_FORMAT_HASH || k->hash_nslots == 0)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "only KDS_FORMAT_HASH inner relations are supported");
            return ERRCODE_DEVICE_ERROR;
        }
        if (ch->kds_offset + k->length > length || k->length >= (1ULL << 32) ||
            KDS_HEAD_LENGTH(k) + 8ULL * k->hash_nslots > k->length)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "inner KDS is truncated or larger than 4GB");
            return ERRCODE_DEVICE_ERROR;
        }
    }
    if (s->d_kmrels)
        (void)hipFree(s->d_kmrels);
    s->d_kmrels = nullptr;
    if (!hip_ok(hipMalloc((void **)&s->d_kmrels, length), eb, "hipMalloc(kmrels)") ||
        !hip_ok(hipMemcpy(s->d_kmrels, kmrels, length, hipMemcpyHostToDevice), eb, "hipMemcpy(kmrels)"))
        return ERRCODE_DEVICE_ERROR;
    uint64_t d_nitems[XP_MAX_DEPTH + 1] = {}, d_nslots[XP_MAX_DEPTH + 1] = {};
    for (uint32_t d = 1; d <= P->nrels; d++)
    {
        const kern_multirels_chunk *ch = &kmrels->chunks[d - 1];
        if (ch->pinned_buffer)
        {
            if ((rc = attach_pinned(s, d, ch->buffer_id, &d_nitems[d], &d_nslots[d], eb)) != 0)
                return rc;
            P->depth[d].inner_kds = (uint64_t)(uintptr_t)s->d_inner_own[d];
        }
        else
        {
            const kern_data_store *k = (const kern_data_store *)((const char *)kmrels + ch->kds_offset);
            d_nitems[d] = k->nitems;
            d_nslots[d] = k->hash_nslots;
            P->depth[d].inner_kds = (uint64_t)(uintptr_t)(s->d_kmrels + ch->kds_offset);
        }
        P->depth[d].left_outer = ch->left_outer ? 1 : 0;
    }
    /* Bloom filters: 16 to 32 bits per item in 64-bit words (a power of two,
     * at most 2^23 words), built from the hash chains just copied */
    if (s->d_bloom)
        (void)hipFree(s->d_bloom);
    s->d_bloom = nullptr;
    uint64_t nwords = 0;
    for (uint32_t d = 1; d <= P->nrels; d++)
    {
        uint64_t w = 1;
        while (w < (d_nitems[d] + 3) / 4 && w < (1ULL << 23))
            w <<= 1;
        s->bloom_off[d] = nwords;
        s->bloom_mask[d] = (uint32_t)(w - 1);
        nwords += w;
    }
    if (nwords > 0)
    {
        /* the filter words, then one counter per depth for the most items
         * sharing a join hash */
        const uint64_t ndup = (P->nrels + 1 + 1) / 2;
        /* then the depth-1 filter folded to at most XQ_SCREEN_LDS_WORDS words */
        const uint64_t w1 = (uint64_t)s->bloom_mask[1] + 1;
        s->bloom_lds_words = (uint32_t)(w1 < XQ_SCREEN_LDS_WORDS ? w1 : XQ_SCREEN_LDS_WORDS);
        s->bloom_lds_off = nwords + ndup;
        if (!hip_ok(hipMalloc((void **)&s->d_bloom, 8 * (nwords + ndup + s->bloom_lds_words)), eb,
                    "hipMalloc(bloom)") ||
            !hip_ok((hipError_t)xpu_launch_fill(s->d_bloom, 0, nwords + ndup, nullptr), eb, "bloom fill"))
            return ERRCODE_DEVICE_ERROR;
        uint32_t *d_dup = (uint32_t *)(s->d_bloom + nwords);
        for (uint32_t d = 1; d <= P->nrels; d++)
        {
            if (!hip_ok((hipError_t)xpu_launch_bloom_build((kern_data_store *)(uintptr_t)P->depth[d].inner_kds,
                                                           d_nslots[d], s->d_bloom + s->bloom_off[d],
                                                           s->bloom_mask[d], d_dup + d, nullptr),
                        eb, "bloom build"))
                return ERRCODE_DEVICE_ERROR;
        }
        if (!hip_ok((hipError_t)xpu_launch_bloom_fold(s->d_bloom + s->bloom_off[1],
                                                      (uint64_t)s->bloom_mask[1] + 1,
                                                      s->d_bloom + s->bloom_lds_off,
                                                      s->bloom_lds_words, nullptr), eb, "bloom fold"))
            return ERRCODE_DEVICE_ERROR;
        uint32_t dup[XP_MAX_DEPTH + 2] = {};
        if (!hip_ok(hipMemcpy(dup, d_dup, 4 * (P->nrels + 1), hipMemcpyDeviceToHost), eb,
                    "bloom build"))
            return ERRCODE_DEVICE_ERROR;
        for (uint32_t d = 1; d <= P->nrels; d++)
            s->maxdup[d] = dup[d];
    }
    if (!hip_ok(hipMemcpy(s->d_plan, P, sizeof(xplan), hipMemcpyHostToDevice), eb, "hipMemcpy(plan)"))
        return ERRCODE_DEVICE_ERROR;
    s->inner_ready = true;
    return 0;
}

/* a scan + group-by block merges its LDS table into kds_final once: below
 * this many rows per block the merges outweigh the rows */
#define XPU_MIN_ROWS_PER_BLOCK  (32ULL << 10)

/* XPUHIP_JIT_AUTO specializes chunks of at least this many rows */
#define XPUHIP_JIT_AUTO_ROWS    1000000ULL

/* device memory for the records of multi-pass joins, per task */
#define XPU_JOIN_REC_BUDGET     (4ULL << 30)

/* the record budget: XPU_JOIN_REC_BUDGET, or the environment's
 * XPUHIP_JOIN_REC_BUDGET (bytes, at least 1 MiB) - smaller budgets split a
 * chunk into more slices without changing results */
static uint64_t
join_rec_budget(void)
{
    const char *e = getenv("XPUHIP_JOIN_REC_BUDGET");
    if (e && *e)
    {
        unsigned long long v = strtoull(e, nullptr, 10);
        if (v >= (1ULL << 20))
            return v;
    }
    return XPU_JOIN_REC_BUDGET;
}

namespace {

/* Sizes of multi-pass join records: the rows per slice (*slice), the
 * per-block capacity and words of each level's records (cap[l], nw[l]: the
 * input of the pass of depth l), and the bytes of the two buffers that the
 * levels alternate between.  False when a slice of one grid-wide round of
 * wave tiles would not fit the budget. */
bool
join_pass_sizes(const xplan *P, const uint32_t *maxdup, uint64_t nrows, uint32_t grid,
                uint32_t nwaves, uint64_t *slice, uint32_t *cap, uint32_t *nw, size_t *bytes)
{
    const uint64_t wrows = 64ULL * XQ_SCREEN_RPT;
    const uint64_t round = (uint64_t)grid * nwaves * wrows;
    /* record words per row of a slice, at the level's most matches per row
     * (a level-l record: the row and the items of depths 1..l-1) */
    uint64_t words[2] = {0, 0}, mult = 1;
    for (uint32_t l = 1; l <= P->nrels; l++)
    {
        uint64_t w = l * mult;
        words[(l - 1) & 1] = w > words[(l - 1) & 1] ? w : words[(l - 1) & 1];
        mult *= maxdup[l] > 1 ? maxdup[l] : 1;
        if (mult > (1ULL << 20))
            return false;
    }
    uint64_t n = join_rec_budget() / (4 * (words[0] + words[1]));
    n = n / round * round;
    if (n == 0)
        return false;
    if (n > nrows)
        n = nrows;
    const uint64_t nwt = (n + wrows - 1) / wrows;
    uint64_t c = ((nwt + (uint64_t)grid * nwaves - 1) / ((uint64_t)grid * nwaves)) * nwaves * wrows;
    bytes[0] = bytes[1] = 0;
    for (uint32_t l = 1; l <= P->nrels; l++)
    {
        if (c > 0xffffffffULL)
            return false;
        cap[l] = (uint32_t)c;
        nw[l] = l;
        size_t b = 4ULL * grid * nw[l] * c;
        bytes[(l - 1) & 1] = b > bytes[(l - 1) & 1] ? b : bytes[(l - 1) & 1];
        c *= maxdup[l] > 1 ? maxdup[l] : 1;
    }
    *slice = n;
    return true;
}

/* code objects of specialized kernels loaded into this process's device */
struct JitFns
{
    hipFunction_t   main = nullptr;         /* scan + group-by or pipeline kernel */
    hipFunction_t   spill = nullptr;        /* spill pass (scan + group-by) */
    hipFunction_t   batch = nullptr;        /* batched pass 1 */
    hipFunction_t   batch_spill = nullptr;  /* batched pass 2 */
    hipFunction_t   dump = nullptr;         /* partial-state dump merge */
    hipFunction_t   count = nullptr;        /* projection count pass */
    hipFunction_t   write = nullptr;        /* projection write pass */
    hipFunction_t   scount = nullptr;       /* scan + projection: count pass */
    hipFunction_t   swrite = nullptr;       /* scan + projection: write pass */
    hipFunction_t   screen = nullptr;       /* multi-pass joins: screen pass */
    hipFunction_t   probe[XP_MAX_DEPTH + 1] = {};  /* multi-pass joins: pass of depth d */
};
std::mutex g_mod_lock;
std::map<const std::vector<char> *, JitFns> g_mods;

const JitFns *
jit_function(xpuhip_session *s, int kind, int param, kern_errorbuf *eb)
{
    std::string log;
    const std::vector<char> *code = xpu_jit_code(&s->comp.plan, kind, param, &log);
    if (!code)
    {
        if (log.size() > 150)
            log.resize(150);
        XERR(eb, ERRCODE_DEVICE_ERROR, ("specialized kernel does not compile: " + log).c_str());
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_mod_lock);
    auto it = g_mods.find(code);
    if (it != g_mods.end())
        return &it->second;
    hipModule_t mod;
    JitFns f;
    if (!hip_ok(hipModuleLoadData(&mod, code->data()), eb, "hipModuleLoadData") ||
        !hip_ok(hipModuleGetFunction(&f.main, mod, "xpu_jit_kernel"), eb, "hipModuleGetFunction"))
        return nullptr;
    if (kind == XJ_SCAN_AGG &&
        (!hip_ok(hipModuleGetFunction(&f.spill, mod, "xpu_jit_spill_kernel"), eb,
                 "hipModuleGetFunction") ||
         !hip_ok(hipModuleGetFunction(&f.batch, mod, "xpu_jit_batch_kernel"), eb,
                 "hipModuleGetFunction") ||
         !hip_ok(hipModuleGetFunction(&f.batch_spill, mod, "xpu_jit_batch_spill_kernel"), eb,
                 "hipModuleGetFunction") ||
         !hip_ok(hipModuleGetFunction(&f.dump, mod, "xpu_jit_dump_kernel"), eb,
                 "hipModuleGetFunction")))
        return nullptr;
    if (kind == XJ_PIPELINE && !param &&
        (!hip_ok(hipModuleGetFunction(&f.count, mod, "xpu_jit_count_kernel"), eb,
                 "hipModuleGetFunction") ||
         !hip_ok(hipModuleGetFunction(&f.write, mod, "xpu_jit_write_kernel"), eb,
                 "hipModuleGetFunction")))
        return nullptr;
    if (kind == XJ_PIPELINE && !param && s->comp.plan.nrels == 0 &&
        (!hip_ok(hipModuleGetFunction(&f.scount, mod, "xpu_jit_scount_kernel"), eb,
                 "hipModuleGetFunction") ||
         !hip_ok(hipModuleGetFunction(&f.swrite, mod, "xpu_jit_swrite_kernel"), eb,
                 "hipModuleGetFunction")))
        return nullptr;
    if (kind == XJ_PIPELINE && param && s->comp.plan.nrels > 0)
    {
        if (!hip_ok(hipModuleGetFunction(&f.screen, mod, "xpu_jit_screen_kernel"), eb,
                    "hipModuleGetFunction"))
            return nullptr;
        for (uint32_t d = 1; d <= s->comp.plan.nrels && d <= XP_MAX_DEPTH; d++)
        {
            char name[64];
            snprintf(name, sizeof(name), "xpu_jit_probe_kernel_%u", d);
            if (!hip_ok(hipModuleGetFunction(&f.probe[d], mod, name), eb, "hipModuleGetFunction"))
                return nullptr;
        }
    }
    return &(g_mods[code] = f);
}

}   /* namespace */

/* ------------------------------------------------------------------
 * Group-by kds_final growth (gpu_service.c:3868-3884 and
 * __expandGpuQueryGroupByBuffer, :2418-2496).  The reference's kernel
 * suspends when a new group finds no room, the service grows the buffer and
 * the kernel resumes.  Here a launch never stops: a group that finds no room
 * in kds_final goes to the session's overflow table (same layout), and after
 * the launch kds_final grows as the reference grows it (by its length, at
 * least 1 GiB, and at least by what the overflow table holds) and the
 * overflow groups are merged into it.  The task then completes.
 * ------------------------------------------------------------------ */
#define XPU_OVF_MIN_BYTES       (256ULL << 20)
#define XPU_OVF_ROWS_BASE       (64ULL << 20)
#define XPU_OVF_BYTES_PER_ROW   256ULL

/* the overflow table's header image: kds_final's with its own length and
 * hash slots, empty */
static std::vector<char>
ovf_head(const xpuhip_session *s, uint64_t length, uint32_t nslots)
{
    std::vector<char> img(s->comp.kds_final_head.begin(), s->comp.kds_final_head.end());
    kern_data_store *h = (kern_data_store *)img.data();
    h->length = length;
    h->usage = 0;
    h->nitems = 0;
    h->hash_nslots = nslots;
    return img;
}

/* Sizes the overflow table for a launch of nrows rows: a quarter of
 * kds_final's length, at least 256 MiB, and no more than 64 MiB plus 256 B
 * per row.  One launch can place that much of new groups beyond kds_final's
 * free space; kds_final then grows by at least 1 GiB before the next one.
 * The table is empty between launches, so it is replaced under the
 * exclusive lock when a longer one is due. */
static int
ensure_ovf(xpuhip_session *s, uint64_t nrows, kern_errorbuf *eb)
{
    uint64_t want = std::max<uint64_t>(s->final_length / 4, XPU_OVF_MIN_BYTES);
    const uint64_t by_rows = XPU_OVF_ROWS_BASE + nrows * XPU_OVF_BYTES_PER_ROW;
    if (want > by_rows)
        want = by_rows;
    want = (want + XPU_L1_CACHELINE_SZ - 1) & ~(uint64_t)(XPU_L1_CACHELINE_SZ - 1);
    {
        std::shared_lock<std::shared_mutex> g(s->final_rw);
        if (s->d_kds_ovf && s->ovf_length >= want)
            return 0;
    }
    std::unique_lock<std::shared_mutex> g(s->final_rw);
    if (s->d_kds_ovf && s->ovf_length >= want)
        return 0;
    if (s->d_kds_ovf)
    {
        (void)hipFree(s->d_kds_ovf);
        s->d_kds_ovf = nullptr;
        s->ovf_length = 0;
    }
    /* chains of about four groups when the table is full */
    const uint32_t nslots = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want / 512, 1024),
                                                         UINT32_MAX / 2);
    std::vector<char> img = ovf_head(s, want, nslots);
    if (!hip_ok(hipMalloc((void **)&s->d_kds_ovf, want), eb, "hipMalloc(group-by overflow table)"))
    {
        s->d_kds_ovf = nullptr;
        return ERRCODE_DEVICE_ERROR;
    }
    if (!hip_ok(hipMemcpy(s->d_kds_ovf, img.data(), img.size(), hipMemcpyHostToDevice), eb, "hipMemcpy") ||
        xpu_launch_fill((char *)s->d_kds_ovf + img.size(), 0, nslots, nullptr) != 0 ||
        !hip_ok(hipDeviceSynchronize(), eb, "overflow table init"))
    {
        (void)hipFree(s->d_kds_ovf);
        s->d_kds_ovf = nullptr;
        return ERRCODE_DEVICE_ERROR;
    }
    s->ovf_length = want;
    return 0;
}

/* bytes of items a KDS_FORMAT_HASH (or FALLBACK) buffer really holds:
 * kds->usage may run past them after a failed reservation, the row index
 * (offsets of the items' tuples, item_head bytes into each item) does not */
static bool
kds_tuple_bytes(const kern_data_store *d_kds, const kern_data_store &h, uint64_t head_len,
                uint64_t *bytes, kern_errorbuf *eb, uint64_t item_head = offsetof(kern_hashitem, t))
{
    *bytes = 0;
    if (h.nitems == 0)
        return true;
    std::vector<uint64_t> ridx(h.nitems);
    if (!hip_ok(hipMemcpy(ridx.data(), (const char *)d_kds + head_len + 8ULL * h.hash_nslots,
                          8ULL * h.nitems, hipMemcpyDeviceToHost), eb, "hipMemcpy(row index)"))
        return false;
    uint64_t most = 0;
    for (uint64_t r : ridx)
        most = r > most ? r : most;
    *bytes = most + item_head;
    return true;
}

/* Grows kds_final and merges the overflow table into it (once, whichever
 * task gets the exclusive lock first; the others find the table empty). */
static int
final_grow_merge(xpuhip_session *s, Scratch *x, kern_errorbuf *eb)
{
    std::unique_lock<std::shared_mutex> g(s->final_rw);
    const uint64_t head_len = s->comp.kds_final_head.size();
    kern_data_store fh, oh;
    if (!hip_ok(hipMemcpy(&fh, s->d_kds_final, sizeof(fh), hipMemcpyDeviceToHost), eb, "hipMemcpy") ||
        !hip_ok(hipMemcpy(&oh, s->d_kds_ovf, sizeof(oh), hipMemcpyDeviceToHost), eb, "hipMemcpy"))
        return ERRCODE_DEVICE_ERROR;
    if (oh.nitems == 0)
        return 0;
    uint64_t uf = 0, uo = 0;
    if (!kds_tuple_bytes(s->d_kds_final, fh, head_len, &uf, eb) ||
        !kds_tuple_bytes(s->d_kds_ovf, oh, head_len, &uo, eb))
        return ERRCODE_DEVICE_ERROR;
    const uint64_t old_len = s->final_length;
    uint64_t new_len = old_len + std::max<uint64_t>(old_len, 1ULL << 30);
    const uint64_t fixed = head_len + 8ULL * fh.hash_nslots;
    const uint64_t tuples = uf + uo;
    const uint64_t need = fixed + tuples + 8 * (tuples / XPU_L1_CACHELINE_SZ) + 4096;
    if (new_len < need)
        new_len = need;
    new_len = (new_len + XPU_L1_CACHELINE_SZ - 1) & ~(uint64_t)(XPU_L1_CACHELINE_SZ - 1);
    kern_data_store *d_new = nullptr;
    if (!hip_ok(hipMalloc((void **)&d_new, new_len), eb, "hipMalloc(kds_final)"))
        return ERRCODE_DEVICE_ERROR;
    /* head, slots and row index stay; the tuples keep their distance from
     * the tail, so chains and row index stay valid */
    fh.length = new_len;
    fh.usage = uf;
    int rc = 0;
    if (!hip_ok(hipMemcpyAsync(d_new, s->d_kds_final, fixed + 8ULL * fh.nitems,
                               hipMemcpyDeviceToDevice, x->stream), eb, "hipMemcpyAsync") ||
        (uf > 0 && !hip_ok(hipMemcpyAsync((char *)d_new + new_len - uf,
                                          (const char *)s->d_kds_final + old_len - uf, uf,
                                          hipMemcpyDeviceToDevice, x->stream), eb, "hipMemcpyAsync")) ||
        !hip_ok(hipMemcpyAsync(d_new, &fh, sizeof(fh), hipMemcpyHostToDevice, x->stream), eb,
                "hipMemcpyAsync"))
        rc = ERRCODE_DEVICE_ERROR;
    if (rc == 0)
    {
        xpu_launch_fill(x->d_ctl, 0, sizeof(xpu_task_ctl) / 8, x->stream);
        if (xpu_launch_final_merge_ovf(s->d_plan, s->comp.plan.nregs, d_new, s->d_kds_ovf,
                                       oh.nitems, x->d_ctl, x->stream) != 0)
            rc = ERRCODE_DEVICE_ERROR;
    }
    /* the overflow table, empty again */
    std::vector<char> img = ovf_head(s, s->ovf_length, oh.hash_nslots);
    if (rc == 0 &&
        (!hip_ok(hipMemcpyAsync(s->d_kds_ovf, img.data(), img.size(), hipMemcpyHostToDevice,
                                x->stream), eb, "hipMemcpyAsync") ||
         xpu_launch_fill((char *)s->d_kds_ovf + img.size(), 0, oh.hash_nslots, x->stream) != 0 ||
         !hip_ok(hipMemcpyAsync(x->h_ctl, x->d_ctl, sizeof(xpu_task_ctl), hipMemcpyDeviceToHost,
                                x->stream), eb, "hipMemcpyAsync") ||
         !hip_ok(hipStreamSynchronize(x->stream), eb, "kds_final merge")))
        rc = ERRCODE_DEVICE_ERROR;
    if (rc == 0 && x->h_ctl->errcode)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, message_of(s, x->h_ctl->errmsg));
        rc = ERRCODE_DEVICE_ERROR;
    }
    if (rc != 0)
    {
        (void)hipStreamSynchronize(x->stream);
        (void)hipFree(d_new);
        return rc;
    }
    (void)hipFree(s->d_kds_final);
    s->d_kds_final = d_new;
    s->final_length = new_len;
    t_last_stats.final_grows++;
    return 0;
}

static int
task_exec_common(xpuhip_session *s, Scratch *x, const kern_data_store *d_src, uint32_t nitems,
                 kern_data_store *d_dst, size_t dst_length, uint64_t *resume_pos,
                 kern_exec_results *res, kern_errorbuf *eb)
{
    const xplan *P = &s->comp.plan;
    bool preagg = s->comp.has_groupby;
    xpuhip_tuning t = s->tuning;
    uint64_t row_begin = resume_pos ? *resume_pos : 0;

    if (!s->inner_ready)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "inner relations are not attached");
        return ERRCODE_DEVICE_ERROR;
    }
    if (!preagg && (!d_dst || dst_length < s->comp.kds_dst_head.size() + 64 ||
                    dst_length % XPU_MAXIMUM_ALIGNOF != 0 || dst_length >= (32ULL << 30)))
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "projection needs a MAXALIGN'ed kds_dst buffer below 32GB");
        return ERRCODE_DEVICE_ERROR;
    }
    if (res)
    {
        memset(res, 0, sizeof(kern_exec_results));
        res->num_rels = P->nrels;
    }
    if (row_begin >= nitems)
    {
        if (resume_pos)
            *resume_pos = 0;
        if (!preagg)
        {
            xpu_launch_fill(x->d_ctl, 0, sizeof(xpu_task_ctl) / 8, x->stream);
            xpu_launch_finalize_dst(d_dst, s->d_kds_dst_head, (uint32_t)s->comp.kds_dst_head.size(),
                                    dst_length, x->d_ctl, x->stream);
            if (!hip_ok(hipStreamSynchronize(x->stream), eb, "hipStreamSynchronize"))
                return ERRCODE_DEVICE_ERROR;
        }
        return 0;
    }
    const bool scan_agg = uses_scan_agg(s);
    uint32_t block = t.block_size ? t.block_size : 256;
    uint32_t rpt = preagg ? (t.rows_per_thread ? t.rows_per_thread : 1) : 1;
    uint64_t nrows = nitems - row_begin;

    /* group-by: a launch holds kds_final shared (growing it takes it
     * exclusively); the groups that find no room in it go to the overflow
     * table */
    const bool use_ovf = preagg && P->nkeys > 0;
    if (use_ovf && ensure_ovf(s, nrows, eb) != 0)
        return ERRCODE_DEVICE_ERROR;
    std::shared_lock<std::shared_mutex> final_hold;
    if (preagg || s->d_kds_fallback)
        final_hold = std::shared_lock<std::shared_mutex>(s->final_rw);

    /* the kernels specialized for the plan, or the interpreter.  Under
     * XPUHIP_JIT_AUTO a join with a group-by is specialized at any size:
     * its interpreter build has written some groups twice under contention
     * (cause not found), its specialized build has not. */
    const bool use_jit = t.jit == XPUHIP_JIT_ON ||
                         (t.jit != XPUHIP_JIT_OFF &&
                          (nrows >= XPUHIP_JIT_AUTO_ROWS || (preagg && P->nrels > 0)));
    hipFunction_t jfn = nullptr, jsfn = nullptr, jcount = nullptr, jwrite = nullptr, jdump = nullptr;
    const JitFns *jf = nullptr;
    if (use_jit)
    {
        jf = scan_agg ? jit_function(s, XJ_SCAN_AGG, (int)(rpt >= 4 ? 4 : rpt), eb)
                      : jit_function(s, XJ_PIPELINE, preagg ? 1 : 0, eb);
        if (jf)
        {
            jfn = jf->main;
            jsfn = jf->spill;
            jcount = jf->count;
            jwrite = jf->write;
            jdump = jf->dump;
        }
        else if (t.jit == XPUHIP_JIT_ON)
            return ERRCODE_DEVICE_ERROR;
    }
    /* the interpreter's kernels take blocks of up to XP_BLOCK_MAX threads */
    if (!jf && block > XP_BLOCK_MAX)
        block = XP_BLOCK_MAX;
    /* a projection without joins runs over wave tiles of 64*XP_PROJ_RPT rows */
    const bool sproj = !preagg && P->nrels == 0 && jf && jf->scount;
    uint32_t rows_per_tile = sproj ? 64 * XP_PROJ_RPT : block * rpt;
    uint32_t ntiles = (uint32_t)((nrows + rows_per_tile - 1) / rows_per_tile);
    /* wave-tile projections: two blocks per CU for chunks of up to 4M rows
     * (1M-row scan: 0.12 ms against 0.14 ms with four), four above
     * (6M rows: 0.38 ms against 0.49 ms; 50M rows: 2.6 ms against 3.9 ms) */
    uint32_t bpc = t.blocks_per_cu ? t.blocks_per_cu : (sproj && nrows > (4ULL << 20) ? 4 : 2);
    uint32_t grid = (uint32_t)g_dev.ncu * bpc;
    if (scan_agg)
    {
        /* a block's LDS table costs an init and a merge into kds_final: give
         * each block at least XPU_MIN_ROWS_PER_BLOCK rows */
        uint64_t by_rows = (nrows + XPU_MIN_ROWS_PER_BLOCK - 1) / XPU_MIN_ROWS_PER_BLOCK;
        if (grid > by_rows)
            grid = (uint32_t)by_rows;
    }
    if (grid > ntiles)
        grid = ntiles;      /* scan_agg: a block tile is its waves' first tiles */
    uint32_t lds_nents = (preagg && P->lds_fast) ? t.lds_nents : 0;
    size_t lds = lds_bytes_for(P, lds_nents);
    if (lds + (scan_agg ? 0 : XP_PIPELINE_STATIC_LDS) > g_dev.lds_per_block)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "group-by LDS table exceeds the LDS of a block");
        return ERRCODE_DEVICE_ERROR;
    }
    if (!preagg && !grow((char **)&x->d_status, &x->status_cap, 16ULL * ntiles, eb))
        return ERRCODE_DEVICE_ERROR;
    if (sproj && !grow((char **)&x->d_block_status, &x->block_status_cap, 16ULL * grid, eb))
        return ERRCODE_DEVICE_ERROR;

    xpu_kparams kp;
    memset(&kp, 0, sizeof(kp));
    kp.plan = s->d_plan;
    kp.kds_src = d_src;
    kp.kds_out = preagg ? s->d_kds_final : d_dst;
    kp.ctl = x->d_ctl;
    kp.tile_status = x->d_status;
    kp.const_pool = s->d_const_pool;
    kp.row_begin = (uint32_t)row_begin;
    kp.row_end = nitems;
    kp.rows_per_tile = rows_per_tile;
    kp.ntiles = ntiles;
    kp.lds_nents = lds_nents;
    kp.params = s->d_params;
    kp.param_nulls = s->comp.param_nulls;
    kp.block_status = sproj ? x->d_block_status : nullptr;
    kp.kds_ovf = use_ovf ? s->d_kds_ovf : nullptr;
    kp.kds_fallback = s->d_kds_fallback;
    for (uint32_t d = 1; d <= P->nrels && d <= XP_MAX_DEPTH; d++)
        kp.inner_kds[d] = P->depth[d].inner_kds;
    if (t.join_filter != XPUHIP_JOIN_FILTER_OFF && s->d_bloom)
    {
        for (uint32_t d = 1; d <= P->nrels && d <= XP_MAX_DEPTH; d++)
        {
            kp.bloom[d] = s->d_bloom + s->bloom_off[d];
            kp.bloom_mask[d] = s->bloom_mask[d];
        }
        kp.prefilter = (preagg && P->nrels > 0) ? 1 : 0;
    }
    /* spill buffer of a scan + group-by: room for a quarter of the rows (at
     * least 1M records); rows beyond it merge into kds_final directly */
    const bool spill = scan_agg && P->spill_ok && lds_nents > 0;
    if (spill)
    {
        uint64_t rec = 8ULL * (1 + P->spill_nregs);
        uint64_t nrec = nrows / 4 > (1ULL << 20) ? nrows / 4 : (1ULL << 20);
        if (!grow((char **)&x->d_spill, &x->spill_cap, nrec * rec, eb))
            return ERRCODE_DEVICE_ERROR;
        kp.spill = x->d_spill;
        kp.spill_cap = x->spill_cap / rec;
        /* LDS entries leave the scan and spill passes as hash-partitioned
         * partial-state records; one block per partition merges each group
         * into kds_final once */
        uint32_t nparts = (uint32_t)g_dev.ncu < XP_DUMP_MAXPARTS ? (uint32_t)g_dev.ncu : XP_DUMP_MAXPARTS;
        uint64_t rw = 1 + P->lds_n64 + (1 + P->lds_n32 + 1) / 2;
        /* one table per block from the scan pass plus the spill pass's
         * dumps of full tables: room for eight tables per block */
        uint64_t per_part = (8ULL * grid * lds_nents) / nparts + 4096;
        if (!grow((char **)&x->d_dump, &x->dump_cap, per_part * nparts * rw * 8, eb) ||
            !grow((char **)&x->d_dump_next, &x->dump_next_cap, 8ULL * XP_DUMP_MAXPARTS, eb))
            return ERRCODE_DEVICE_ERROR;
        kp.dump = x->d_dump;
        kp.dump_cap = x->dump_cap / (nparts * rw * 8);
        kp.dump_next = x->d_dump_next;
        kp.dump_nparts = nparts;
    }

    /* a group-by over joins with the specialized kernels: a screen pass
     * and one pass per depth, over slices of the rows sized to the record
     * budget */
    bool passes = false;
    uint64_t slice = 0;
    uint32_t rcap[XP_MAX_DEPTH + 2] = {}, rnw[XP_MAX_DEPTH + 2] = {};
    if (preagg && !scan_agg && P->nrels > 0 && use_jit && t.join_passes != XPUHIP_JOIN_PASSES_OFF)
    {
        size_t bytes[2];
        passes = join_pass_sizes(P, s->maxdup, nrows, grid, XQ_SCREEN_BLOCK / 64, &slice, rcap, rnw,
                                 bytes) &&
                 (t.join_passes == XPUHIP_JOIN_PASSES_ON || slice >= nrows || slice >= (1ULL << 20));
        if (passes &&
            (!grow((char **)&x->d_rec[0], &x->rec_cap[0], bytes[0], eb) ||
             !grow((char **)&x->d_rec[1], &x->rec_cap[1], bytes[1], eb) ||
             !grow((char **)&x->d_rec_cnt, &x->rec_cnt_cap, 8ULL * grid, eb)))
            return ERRCODE_DEVICE_ERROR;
        kp.screened = (kp.prefilter && kp.bloom[1] && !P->depth[1].left_outer) ? 1 : 0;
        if (kp.screened && s->bloom_lds_words > 0)
        {
            kp.bloom_lds = s->d_bloom + s->bloom_lds_off;
            kp.bloom_lds_words = s->bloom_lds_words;
        }
    }

    int rc = 0;
    uint32_t nslices = 0;
    xpu_launch_fill(x->d_ctl, 0, sizeof(xpu_task_ctl) / 8, x->stream);
    if (kp.dump)
        xpu_launch_fill(kp.dump_next, 0, kp.dump_nparts, x->stream);
    if (!preagg)
    {
        /* kds_dst header first, so the kernel sees the capacity */
        xpu_launch_finalize_dst(d_dst, s->d_kds_dst_head, (uint32_t)s->comp.kds_dst_head.size(),
                                dst_length, x->d_ctl, x->stream);
        xpu_launch_fill(x->d_status, 0, 2ULL * ntiles, x->stream);
    }
    if (passes && !(jf && jf->screen))
        passes = false;
    (void)hipEventRecord(x->ev0, x->stream);
    if (!preagg)
    {
        /* projection: count pass (tile totals), tile scan, write pass; no
         * tile waits on another */
        if (sproj)
        {
            void *args[] = {&kp};
            rc = (int)hipModuleLaunchKernel(jf->scount, grid, 1, 1, block, 1, 1, 0, x->stream,
                                            args, nullptr);
            if (!rc)
                rc = xpu_launch_tile_scan_blocks(&kp, grid, x->stream);
            if (!rc)
                rc = (int)hipModuleLaunchKernel(jf->swrite, grid, 1, 1, block, 1, 1, 0, x->stream,
                                                args, nullptr);
        }
        else if (jcount && jwrite)
        {
            void *args[] = {&kp};
            rc = (int)hipModuleLaunchKernel(jcount, grid, 1, 1, block, 1, 1, 0, x->stream,
                                            args, nullptr);
            if (!rc)
                rc = xpu_launch_tile_scan(&kp, x->stream);
            if (!rc)
                rc = (int)hipModuleLaunchKernel(jwrite, grid, 1, 1, block, 1, 1, 0, x->stream,
                                                args, nullptr);
        }
        else
            rc = xpu_launch_projection(&kp, P->nregs, grid, block, x->stream);
    }
    else if (passes)
    {
        /* per slice: the screen pass into level-1 records (buffer 0), then
         * the pass of each depth from level d (buffer (d-1)&1) to level d+1 */
        for (uint64_t sb = row_begin; rc == 0 && sb < nitems; sb += slice)
        {
            nslices++;
            xpu_kparams q = kp;
            void *args[] = {&q};
            q.row_begin = (uint32_t)sb;
            q.row_end = (uint32_t)(sb + slice < nitems ? sb + slice : nitems);
            q.pass_depth = 0;
            q.rec_out = x->d_rec[0];
            q.rec_out_cnt = x->d_rec_cnt;
            q.rec_out_cap = rcap[1];
            rc = (int)hipModuleLaunchKernel(jf->screen, grid, 1, 1, XQ_SCREEN_BLOCK, 1, 1,
                                            8U * kp.bloom_lds_words, x->stream, args, nullptr);
            for (uint32_t d = 1; rc == 0 && d <= P->nrels; d++)
            {
                const bool last = (d == P->nrels);
                q.pass_depth = d;
                q.rec_in = x->d_rec[(d - 1) & 1];
                q.rec_in_cnt = x->d_rec_cnt + (size_t)((d - 1) & 1) * grid;
                q.rec_in_cap = rcap[d];
                q.rec_out = last ? nullptr : x->d_rec[d & 1];
                q.rec_out_cnt = last ? nullptr : x->d_rec_cnt + (size_t)(d & 1) * grid;
                q.rec_out_cap = last ? 0 : rcap[d + 1];
                rc = (int)hipModuleLaunchKernel(jf->probe[d], grid, 1, 1, block, 1, 1,
                                                (unsigned)(last ? lds : 0), x->stream, args,
                                                nullptr);
            }
        }
    }
    else if (jfn)
    {
        void *args[] = {&kp};
        rc = (int)hipModuleLaunchKernel(jfn, grid, 1, 1, block, 1, 1, (unsigned)lds, x->stream,
                                        args, nullptr);
    }
    else if (scan_agg)
        rc = xpu_launch_scan_agg(&kp, P->nregs, (int)rpt, grid, block, lds, x->stream);
    else
        rc = xpu_launch_pipeline(&kp, P->nregs, grid, block, lds, preagg, x->stream);
    /* second pass over the spilled rows, stream-ordered after the first (it
     * reads the spill count and kds_final's group count on the device) */
    if (rc == 0 && spill)
    {
        (void)hipEventRecord(x->evm, x->stream);
        uint32_t sgrid = (uint32_t)g_dev.ncu * (t.blocks_per_cu ? t.blocks_per_cu : 1);
        /* the spill and dump passes run blocks of at most XP_BLOCK_MAX threads */
        const uint32_t sblock = block < XP_BLOCK_MAX ? block : XP_BLOCK_MAX;
        if (jfn && jsfn)
        {
            void *args[] = {&kp};
            rc = (int)hipModuleLaunchKernel(jsfn, sgrid, 1, 1, sblock, 1, 1, (unsigned)lds,
                                            x->stream, args, nullptr);
        }
        else
            rc = xpu_launch_spill_agg(&kp, P->nregs, sgrid, sblock, lds, x->stream);
        /* the partial-state dumps of both passes, one block per partition */
        if (rc == 0 && kp.dump)
        {
            uint32_t dgrid = kp.dump_nparts;
            if (jfn && jdump)
            {
                void *args[] = {&kp};
                rc = (int)hipModuleLaunchKernel(jdump, dgrid, 1, 1, sblock, 1, 1, (unsigned)lds,
                                                x->stream, args, nullptr);
            }
            else
                rc = xpu_launch_dump_merge(&kp, P->nregs, dgrid, sblock, lds, x->stream);
        }
    }
    (void)hipEventRecord(x->ev1, x->stream);
    if (rc != 0)
    {
        hip_ok((hipError_t)rc, eb, "pipeline kernel launch");
        return ERRCODE_DEVICE_ERROR;
    }
    if (!preagg)
        xpu_launch_finalize_dst(d_dst, s->d_kds_dst_head, (uint32_t)s->comp.kds_dst_head.size(),
                                dst_length, x->d_ctl, x->stream);
    (void)hipMemcpyAsync(x->h_ctl, x->d_ctl, sizeof(xpu_task_ctl), hipMemcpyDeviceToHost, x->stream);
    if (!hip_ok(hipStreamSynchronize(x->stream), eb, "pipeline kernel"))
        return ERRCODE_DEVICE_ERROR;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, x->ev0, x->ev1);
    float ms1 = ms;
    if (spill)
        (void)hipEventElapsedTime(&ms1, x->ev0, x->evm);
    t_last_kernel_ms = ms;

    const xpu_task_ctl *c = x->h_ctl;
    t_last_stats.kernel_ms = ms;
    t_last_stats.pass1_ms = ms1;
    t_last_stats.spill_ms = ms - ms1;
    t_last_stats.grid = grid;
    t_last_stats.block = block;
    t_last_stats.ntiles = ntiles;
    t_last_stats.lds_nents = lds_nents;
    t_last_stats.lds_bytes = lds;
    t_last_stats.rows = nrows;
    t_last_stats.lds_miss = c->lds_miss;
    t_last_stats.global_merges = c->global_merges;
    t_last_stats.spill_rows = c->spill_rows;
    t_last_stats.jit = jf ? 1 : 0;
    t_last_stats.join_slices = nslices;
    t_last_stats.final_grows = 0;
    t_last_stats.heap_ms = 0.0;
    t_last_stats.heap_rows = 0;
    if (res)
    {
        res->nitems_raw = (uint32_t)nrows;
        res->nitems_in = c->nitems_in;
        res->nitems_out = c->nitems_out;
        for (uint32_t d = 1; d <= P->nrels && d <= 8; d++)
            res->stats[d - 1].nitems_out = c->depth_out[d];
    }
    if (c->errcode && c->errcode != ERRCODE_SUSPEND_NO_SPACE)
    {
        XERR(eb, c->errcode, message_of(s, c->errmsg));
        return (int)c->errcode;
    }
    if (c->errcode == ERRCODE_SUSPEND_NO_SPACE)
    {
        /* neither kds_final nor the overflow table had room for a group:
         * its contribution is lost */
        XERR(eb, ERRCODE_DEVICE_ERROR, message_of(s, c->errmsg));
        return ERRCODE_DEVICE_ERROR;
    }
    if (preagg && c->final_ovf)
    {
        /* kds_final grows and takes the overflow groups: the task completes */
        final_hold.unlock();
        int grc = final_grow_merge(s, x, eb);
        if (grc != 0)
            return grc;
    }
    if (!preagg && c->resume_tile != 0)
    {
        /* resume_tile holds ~(first tile that did not fit) */
        uint64_t next = row_begin + (uint64_t)(0xffffffffU - c->resume_tile) * rows_per_tile;
        if (resume_pos)
            *resume_pos = next;
        XERR(eb, ERRCODE_SUSPEND_NO_SPACE, "GpuProjection - no space to write");
        return ERRCODE_SUSPEND_NO_SPACE;
    }
    if (resume_pos)
        *resume_pos = 0;
    return 0;
}

/* projected rows staged per round before they go to a pinned buffer */
#define XPU_PINNED_STAGING      (64ULL << 20)

/* A projection into the session's pinned hash buffer: the chunk's rows go
 * through the projection into a ROW staging buffer (resuming while it
 * fills), and each round's rows are appended to the buffer as hash items
 * (cuda_gpujoin.cu:898-916).  Nothing returns to the host. */
static int
task_exec_pinned(xpuhip_session *s, Scratch *x, const kern_data_store *d_src, uint32_t nitems,
                 kern_exec_results *res, kern_errorbuf *eb)
{
    if (!grow(&x->d_dst, &x->dst_cap, XPU_PINNED_STAGING + 64, eb))
        return ERRCODE_DEVICE_ERROR;
    uint64_t *d_base = (uint64_t *)(x->d_dst + XPU_PINNED_STAGING);
    kern_exec_results sum;
    memset(&sum, 0, sizeof(sum));
    sum.num_rels = s->comp.plan.nrels;
    uint64_t pos = 0;
    for (;;)
    {
        kern_exec_results part;
        int rc = task_exec_common(s, x, d_src, nitems, (kern_data_store *)x->d_dst,
                                  XPU_PINNED_STAGING, &pos, &part, eb);
        if (rc != 0 && rc != ERRCODE_SUSPEND_NO_SPACE)
            return rc;
        const uint32_t nrows = (uint32_t)x->h_ctl->out_nitems;
        if (!hip_ok((hipError_t)xpu_launch_hash_append(s->d_kds_final, (const kern_data_store *)x->d_dst,
                                                      d_base, nrows, x->stream),
                    eb, "pinned buffer append") ||
            !hip_ok(hipMemcpyAsync(x->h_hstat, d_base, 16, hipMemcpyDeviceToHost, x->stream), eb,
                    "hipMemcpyAsync") ||
            !hip_ok(hipStreamSynchronize(x->stream), eb, "pinned buffer append"))
            return ERRCODE_DEVICE_ERROR;
        const uint64_t *base = (const uint64_t *)x->h_hstat;
        if (base[0] == ~0ULL)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, base[1] == 1 ? "pinned inner buffer is full"
                                                        : "pinned inner buffer lock timed out");
            return ERRCODE_DEVICE_ERROR;
        }
        sum.nitems_raw = part.nitems_raw > sum.nitems_raw ? part.nitems_raw : sum.nitems_raw;
        sum.nitems_in += part.nitems_in;
        sum.nitems_out += part.nitems_out;
        for (uint32_t d = 0; d < s->comp.plan.nrels && d < 8; d++)
            sum.stats[d].nitems_out += part.stats[d].nitems_out;
        if (rc == 0)
            break;
    }
    sum.nitems_raw = nitems;
    if (res)
        *res = sum;
    return 0;
}

/* ------------------------------------------------------------------
 * heap-page (KDS_FORMAT_BLOCK) and row (KDS_FORMAT_ROW / HASH) chunks
 *
 * The chunk at d_src (host copy of its header: head) is read into column
 * tiles in x->d_tiles (xpu_heap.hip): a count pass and a scan of the per-unit
 * counts, a read-back of the total (it sizes the tiles and the pipeline's
 * launch), then the deform pass.  The tiles form a KDS_FORMAT_ARROW chunk of
 * *rows rows holding the columns the session's programs load.
 * ------------------------------------------------------------------ */
static uint64_t
align64(uint64_t v)
{
    return (v + 63) & ~63ULL;
}

static int
heap_tiles(xpuhip_session *s, Scratch *x, const kern_data_store *d_src, const kern_data_store *head,
           uint32_t *rows, double *ms, kern_errorbuf *eb)
{
    const xpu_compiled *C = &s->comp;
    const char fmt = head->format;
    char msg[160];
    *rows = 0;
    *ms = 0.0;
    if (fmt != KDS_FORMAT_BLOCK && fmt != KDS_FORMAT_ROW && fmt != KDS_FORMAT_HASH)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "kds_src is not a KDS_FORMAT_BLOCK, ROW or HASH chunk");
        return ERRCODE_DEVICE_ERROR;
    }
    const uint64_t hl = KDS_HEAD_LENGTH(head);
    if (head->nr_colmeta < head->ncols || head->length < hl || head->length >= (1ULL << 32))
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "heap / row chunk header is malformed or 4GB or larger");
        return ERRCODE_DEVICE_ERROR;
    }
    const uint32_t natts = (uint32_t)C->src_cols.size();
    if (natts > head->ncols)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "kds_src has fewer columns than the plan loads");
        return ERRCODE_DEVICE_ERROR;
    }
    for (uint32_t j = 0; j < natts; j++)
    {
        const kern_colmeta *cm = &head->colmeta[j];
        const xpu_src_col &sc = C->src_cols[j];
        int a = cm->attalign;
        bool ok = (cm->attlen > 0 || cm->attlen == -1) && (a == 1 || a == 2 || a == 4 || a == 8);
        if (sc.kind == XH_FIXED)
            ok = ok && cm->attlen == (int)sc.width;
        else if (sc.kind == XH_BOOL)
            ok = ok && cm->attlen == 1;
        else if (sc.kind == XH_VAR)
            ok = ok && cm->attlen == -1;
        if (!ok)
        {
            snprintf(msg, sizeof(msg), "kds_src column %u (attlen %d, attalign %d) cannot be read "
                     "as the plan loads it", j + 1, (int)cm->attlen, a);
            XERR(eb, ERRCODE_DEVICE_ERROR, msg);
            return ERRCODE_DEVICE_ERROR;
        }
    }
    xheap_params hp;
    memset(&hp, 0, sizeof(hp));
    hp.src = d_src;
    hp.src_length = head->length;
    hp.format = (uint32_t)(uint8_t)fmt;
    hp.nitems = head->nitems;
    hp.src_ncols = head->ncols;
    hp.natts_walk = natts;
    hp.xids = s->d_xids;
    hp.nxids = (uint32_t)C->xids.size();
    if (fmt == KDS_FORMAT_BLOCK)
    {
        if (head->block_offset % 8 != 0 || head->block_offset < hl + 4ULL * head->nitems ||
            head->block_offset + (uint64_t)XPU_BLCKSZ * head->nitems > head->length)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "KDS_FORMAT_BLOCK pages out of the chunk");
            return ERRCODE_DEVICE_ERROR;
        }
        hp.block_offset = head->block_offset;
        hp.nunits = head->nitems;
    }
    else
    {
        hp.rowindex_off = hl + 8ULL * head->hash_nslots;
        if (hp.rowindex_off + 8ULL * head->nitems > head->length)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "row index out of the chunk");
            return ERRCODE_DEVICE_ERROR;
        }
        hp.nunits = (head->nitems + 63) / 64;
    }
    /* status, column table, unit counts */
    const size_t cols_off = 16;
    const size_t units_off = (cols_off + sizeof(xheap_col) * natts + 15) & ~(size_t)15;
    /* the unit scan reads and writes whole groups of 4 units past nunits */
    if (!grow(&x->d_hmeta, &x->hmeta_cap, units_off + 4ULL * (hp.nunits + 8), eb))
        return ERRCODE_DEVICE_ERROR;
    hp.st = (xheap_status *)x->d_hmeta;
    hp.cols = (const xheap_col *)(x->d_hmeta + cols_off);
    hp.units = (uint32_t *)(x->d_hmeta + units_off);

    (void)hipEventRecord(x->evh0, x->stream);
    xpu_launch_fill(hp.st, 0, sizeof(xheap_status) / 8, x->stream);
    int rc = xpu_launch_heap_count(&hp, x->stream);
    (void)hipEventRecord(x->evh1, x->stream);
    if (rc != 0)
    {
        hip_ok((hipError_t)rc, eb, "heap count launch");
        return ERRCODE_DEVICE_ERROR;
    }
    (void)hipMemcpyAsync(x->h_hstat, hp.st, sizeof(xheap_status), hipMemcpyDeviceToHost, x->stream);
    if (!hip_ok(hipStreamSynchronize(x->stream), eb, "heap count pass"))
        return ERRCODE_DEVICE_ERROR;
    float ms0 = 0.0f;
    (void)hipEventElapsedTime(&ms0, x->evh0, x->evh1);
    *ms = ms0;
    if (x->h_hstat->errcode)
    {
        XERR(eb, x->h_hstat->errcode, message_of(s, x->h_hstat->errmsg));
        return (int)x->h_hstat->errcode;
    }
    const uint32_t total = x->h_hstat->total;

    /* tile layout: header, validity bitmaps and boolean values (zeroed in
     * one fill), then values and text pairs */
    std::vector<char> &H = x->h_tiles_head;
    H.assign((const char *)head, (const char *)head + hl);
    kern_data_store *th = (kern_data_store *)H.data();
    th->format = KDS_FORMAT_ARROW;
    th->nitems = total;
    th->usage = 0;
    th->hash_nslots = 0;
    th->block_offset = 0;
    th->block_nloaded = 0;
    for (uint32_t j = 0; j < th->ncols; j++)
    {
        kern_colmeta *cm = &th->colmeta[j];
        cm->kds_format = KDS_FORMAT_ARROW;
        cm->nullmap_offset = cm->nullmap_length = 0;
        cm->values_offset = cm->values_length = 0;
        cm->extra_offset = cm->extra_length = 0;
    }
    const uint64_t bm_bytes = 4ULL * ((total + 31) / 32 > 0 ? (total + 31) / 32 : 1);
    uint64_t pos = align64(hl);
    const uint64_t bm_begin = pos;
    for (uint32_t j = 0; j < natts; j++)
    {
        const xpu_src_col &sc = C->src_cols[j];
        if (!sc.kind)
            continue;
        kern_colmeta *cm = &th->colmeta[j];
        cm->nullmap_offset = pos;
        cm->nullmap_length = bm_bytes;
        pos += align64(bm_bytes);
        if (sc.kind == XH_BOOL)
        {
            cm->values_offset = pos;
            cm->values_length = bm_bytes;
            cm->attopts.tag = ArrowType__Bool;
            pos += align64(bm_bytes);
        }
    }
    const uint64_t bm_end = pos;
    for (uint32_t j = 0; j < natts; j++)
    {
        const xpu_src_col &sc = C->src_cols[j];
        kern_colmeta *cm = &th->colmeta[j];
        if (sc.kind == XH_FIXED)
        {
            cm->values_offset = pos;
            cm->values_length = (uint64_t)total * sc.width;
            cm->attopts.unitsz = sc.width;
            pos += align64(cm->values_length);
        }
        else if (sc.kind == XH_VAR)
        {
            cm->values_offset = pos;
            cm->values_length = 8ULL * total;
            cm->attopts.tag = XARROW_TAG_HEAPREF;
            cm->attopts.unitsz = 8;
            cm->extra_length = head->length;
            pos += align64(cm->values_length);
        }
    }
    th->length = pos;
    if (!grow(&x->d_tiles, &x->tiles_cap, pos, eb))
        return ERRCODE_DEVICE_ERROR;
    const uint64_t src_rel = (uint64_t)((uintptr_t)d_src - (uintptr_t)x->d_tiles);
    x->h_hcols.assign(natts, xheap_col{});
    for (uint32_t j = 0; j < natts; j++)
    {
        const kern_colmeta *sm = &head->colmeta[j];
        kern_colmeta *cm = &th->colmeta[j];
        xheap_col &c = x->h_hcols[j];
        c.attlen = sm->attlen;
        c.attalign = sm->attalign;
        c.kind = C->src_cols[j].kind;
        c.width = C->src_cols[j].width;
        c.values_off = cm->values_offset;
        c.nullmap_off = cm->nullmap_offset;
        if (c.kind == XH_VAR)
            cm->extra_offset = src_rel;
    }
    hp.dst = x->d_tiles;
    hp.capacity = total;
    (void)hipEventRecord(x->evh0, x->stream);
    if (!hip_ok(hipMemcpyAsync(x->d_tiles, H.data(), H.size(), hipMemcpyHostToDevice, x->stream),
                eb, "hipMemcpyAsync(tiles)") ||
        (natts > 0 &&
         !hip_ok(hipMemcpyAsync((void *)hp.cols, x->h_hcols.data(), sizeof(xheap_col) * natts,
                                hipMemcpyHostToDevice, x->stream), eb, "hipMemcpyAsync(tiles)")))
        return ERRCODE_DEVICE_ERROR;
    xpu_launch_fill(x->d_tiles + bm_begin, 0, (bm_end - bm_begin) / 8, x->stream);
    rc = xpu_launch_heap_deform(&hp, g_dev.ncu, x->stream);
    (void)hipEventRecord(x->evh1, x->stream);
    if (rc != 0)
    {
        hip_ok((hipError_t)rc, eb, "heap deform launch");
        return ERRCODE_DEVICE_ERROR;
    }
    *rows = total;
    return 0;
}

/* a heap-page / row chunk at d_src: tiles, then the pipeline over them */
static int
task_exec_tuples(xpuhip_session *s, Scratch *x, const kern_data_store *d_src,
                 const kern_data_store *head, kern_data_store *d_dst, size_t dst_length,
                 uint64_t *resume_pos, kern_exec_results *res, kern_errorbuf *eb)
{
    uint32_t rows = 0;
    double ms0 = 0.0;
    int rc = heap_tiles(s, x, d_src, head, &rows, &ms0, eb);
    if (rc != 0)
        return rc;
    if (s->comp.pinned_hash)
        rc = task_exec_pinned(s, x, (const kern_data_store *)x->d_tiles, rows, res, eb);
    else
        rc = task_exec_common(s, x, (const kern_data_store *)x->d_tiles, rows, d_dst, dst_length,
                              resume_pos, res, eb);
    /* errors of the deform pass come first: the tiles are not valid then */
    (void)hipMemcpyAsync(x->h_hstat, x->d_hmeta, sizeof(xheap_status), hipMemcpyDeviceToHost,
                         x->stream);
    if (!hip_ok(hipStreamSynchronize(x->stream), eb, "heap deform pass"))
        return ERRCODE_DEVICE_ERROR;
    if (x->h_hstat->errcode)
    {
        XERR(eb, x->h_hstat->errcode, message_of(s, x->h_hstat->errmsg));
        return (int)x->h_hstat->errcode;
    }
    float ms1 = 0.0f;
    (void)hipEventElapsedTime(&ms1, x->evh0, x->evh1);
    t_last_stats.heap_ms = ms0 + ms1;
    t_last_stats.heap_rows = rows;
    return rc;
}

extern "C" int
xpuhip_task_exec_device_tuples(xpuhip_session *s, const kern_data_store *d_kds_src,
                               const kern_data_store *kds_head, kern_data_store *d_kds_dst,
                               size_t dst_length, uint64_t *resume_pos, kern_exec_results *res,
                               kern_errorbuf *eb)
{
    int rc;
    if (!s || !d_kds_src || !kds_head)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no session or chunk");
        return ERRCODE_DEVICE_ERROR;
    }
    if ((rc = ensure_device(eb)) != 0)
        return rc;
    Scratch *x = scratch_get(s, eb);
    if (!x)
        return ERRCODE_DEVICE_ERROR;
    rc = task_exec_tuples(s, x, d_kds_src, kds_head, d_kds_dst, dst_length, resume_pos, res, eb);
    scratch_put(s, x);
    return rc;
}

extern "C" int
xpuhip_task_exec_device(xpuhip_session *s, const kern_data_store *d_kds_src, uint32_t src_nitems,
                        kern_data_store *d_kds_dst, size_t dst_length, uint64_t *resume_pos,
                        kern_exec_results *res, kern_errorbuf *eb)
{
    int rc;
    if (!s)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no session");
        return ERRCODE_DEVICE_ERROR;
    }
    if ((rc = ensure_device(eb)) != 0)
        return rc;
    Scratch *x = scratch_get(s, eb);
    if (!x)
        return ERRCODE_DEVICE_ERROR;
    if (s->comp.pinned_hash)
        rc = task_exec_pinned(s, x, d_kds_src, src_nitems, res, eb);
    else
        rc = task_exec_common(s, x, d_kds_src, src_nitems, d_kds_dst, dst_length, resume_pos, res, eb);
    scratch_put(s, x);
    return rc;
}

/* XpuTaskExec on a device-resident Arrow chunk that allocates nothing: for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops).  Scan + projection and scan + group-by sessions
 * without joins; the ahead-of-time kernels, on the null stream, with the
 * task scratch in the file-scope array above.  Calls from threads of one
 * process serialise on a lock (they share the device's array, which keeps
 * the last call's contents; everything a call reads is cleared first). */
namespace {
std::mutex g_noalloc_lock;
}

extern "C" int
xpuhip_task_exec_device_noalloc(xpuhip_session *s, const kern_data_store *d_src, uint32_t nitems,
                                kern_data_store *d_dst, size_t dst_length, uint64_t *resume_pos,
                                kern_exec_results *res, kern_errorbuf *eb)
{
    if (!s)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no session");
        return ERRCODE_DEVICE_ERROR;
    }
    if (ensure_device(eb) != 0)
        return ERRCODE_DEVICE_ERROR;
    char *base = g_scratch_ptr[g_dev.id];
    const xplan *P = &s->comp.plan;
    const bool preagg = s->comp.has_groupby;
    if (!base)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no-alloc scratch has no device address");
        return ERRCODE_DEVICE_ERROR;
    }
    if (P->nrels > 0 || s->comp.pinned_hash || (preagg && !uses_scan_agg(s)))
    {
        XERR(eb, ERRCODE_DEVICE_ERROR,
             "the no-alloc entry point runs scan + projection and scan + group-by sessions only");
        return ERRCODE_DEVICE_ERROR;
    }
    if (!preagg && (!d_dst || dst_length < s->comp.kds_dst_head.size() + 64 ||
                    dst_length % XPU_MAXIMUM_ALIGNOF != 0 || dst_length >= (32ULL << 30)))
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "projection needs a MAXALIGN'ed kds_dst buffer below 32GB");
        return ERRCODE_DEVICE_ERROR;
    }
    uint64_t row_begin = resume_pos ? *resume_pos : 0;
    if (res)
    {
        memset(res, 0, sizeof(kern_exec_results));
        res->num_rels = 0;
    }
    xpuhip_tuning t = s->tuning;
    uint32_t block = t.block_size ? t.block_size : 256;
    if (block > XP_BLOCK_MAX)
        block = XP_BLOCK_MAX;
    uint32_t rpt = preagg ? (t.rows_per_thread ? t.rows_per_thread : 1) : 1;
    rpt = rpt >= 4 ? 4 : rpt >= 2 ? 2 : 1;
    const uint64_t nrows = row_begin < nitems ? nitems - row_begin : 0;
    const uint32_t rows_per_tile = block * rpt;
    const uint64_t ntiles = (nrows + rows_per_tile - 1) / rows_per_tile;
    /* scratch layout: control block, then per task kind */
    const size_t ctl_bytes = (sizeof(xpu_task_ctl) + 255) & ~(size_t)255;
    const size_t avail = XPU_NOALLOC_SCRATCH_BYTES - ctl_bytes;
    if (!preagg && 16ULL * ntiles > avail)
    {
        char msg[160];
        snprintf(msg, sizeof(msg), "chunk of %llu rows exceeds the no-alloc scratch (at most %llu rows)",
                 (unsigned long long)nrows, (unsigned long long)(avail / 16 * rows_per_tile));
        XERR(eb, ERRCODE_DEVICE_ERROR, msg);
        return ERRCODE_DEVICE_ERROR;
    }
    std::lock_guard<std::mutex> g(g_noalloc_lock);
    /* kds_final cannot grow here (that allocates): a group-by whose groups
     * do not fit fails with 'E' */
    std::shared_lock<std::shared_mutex> final_hold(s->final_rw);
    xpu_task_ctl *d_ctl = (xpu_task_ctl *)base;
    char *area = base + ctl_bytes;
    if (nrows == 0)
    {
        if (resume_pos)
            *resume_pos = 0;
        if (!preagg)
        {
            xpu_launch_fill(d_ctl, 0, sizeof(xpu_task_ctl) / 8, nullptr);
            xpu_launch_finalize_dst(d_dst, s->d_kds_dst_head, (uint32_t)s->comp.kds_dst_head.size(),
                                    dst_length, d_ctl, nullptr);
            if (!hip_ok(hipStreamSynchronize(nullptr), eb, "hipStreamSynchronize"))
                return ERRCODE_DEVICE_ERROR;
        }
        return 0;
    }
    const bool scan_agg = preagg;
    uint32_t bpc = t.blocks_per_cu ? t.blocks_per_cu : 2;
    uint64_t grid = (uint64_t)g_dev.ncu * bpc;
    if (scan_agg)
    {
        uint64_t by_rows = (nrows + XPU_MIN_ROWS_PER_BLOCK - 1) / XPU_MIN_ROWS_PER_BLOCK;
        grid = grid < by_rows ? grid : by_rows;
    }
    grid = grid < ntiles ? grid : ntiles;
    uint32_t lds_nents = (preagg && P->lds_fast) ? t.lds_nents : 0;
    size_t lds = lds_bytes_for(P, lds_nents);
    if (lds + (scan_agg ? 0 : XP_PIPELINE_STATIC_LDS) > g_dev.lds_per_block)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "group-by LDS table exceeds the LDS of a block");
        return ERRCODE_DEVICE_ERROR;
    }
    xpu_kparams kp;
    memset(&kp, 0, sizeof(kp));
    kp.plan = s->d_plan;
    kp.kds_src = d_src;
    kp.kds_out = preagg ? s->d_kds_final : d_dst;
    kp.ctl = d_ctl;
    kp.const_pool = s->d_const_pool;
    kp.row_begin = (uint32_t)row_begin;
    kp.row_end = nitems;
    kp.rows_per_tile = rows_per_tile;
    kp.ntiles = (uint32_t)ntiles;
    kp.lds_nents = lds_nents;
    kp.params = s->d_params;
    kp.param_nulls = s->comp.param_nulls;
    if (!preagg)
        kp.tile_status = (uint64_t *)area;
    /* group-by: dump counters, then the dump records and the spill records
     * share the rest; rows or entries beyond either merge into kds_final */
    const bool spill = scan_agg && P->spill_ok && lds_nents > 0;
    if (spill)
    {
        const uint32_t nparts = (uint32_t)g_dev.ncu < XP_DUMP_MAXPARTS ? (uint32_t)g_dev.ncu
                                                                      : XP_DUMP_MAXPARTS;
        const uint64_t rw = 1 + P->lds_n64 + (1 + P->lds_n32 + 1) / 2;
        const uint64_t rec = 8ULL * (1 + P->spill_nregs);
        const size_t next_bytes = (8ULL * XP_DUMP_MAXPARTS + 255) & ~(size_t)255;
        const size_t rest = avail - next_bytes;
        const uint64_t dump_cap = (rest / 2) / (nparts * rw * 8);
        const size_t dump_bytes = ((size_t)dump_cap * nparts * rw * 8 + 255) & ~(size_t)255;
        kp.dump_next = (uint64_t *)area;
        kp.dump = (uint64_t *)(area + next_bytes);
        kp.dump_cap = dump_cap;
        kp.dump_nparts = nparts;
        kp.spill = (uint64_t *)(area + next_bytes + dump_bytes);
        kp.spill_cap = (rest - dump_bytes) / rec;
    }
    int rc = xpu_launch_fill(d_ctl, 0, sizeof(xpu_task_ctl) / 8, nullptr);
    if (rc == 0 && kp.dump)
        rc = xpu_launch_fill(kp.dump_next, 0, kp.dump_nparts, nullptr);
    if (rc == 0 && !preagg)
    {
        rc = xpu_launch_finalize_dst(d_dst, s->d_kds_dst_head, (uint32_t)s->comp.kds_dst_head.size(),
                                     dst_length, d_ctl, nullptr);
        if (rc == 0)
            rc = xpu_launch_fill(kp.tile_status, 0, 2ULL * ntiles, nullptr);
        if (rc == 0)
            rc = xpu_launch_projection(&kp, P->nregs, (int)grid, (int)block, nullptr);
    }
    else if (rc == 0)
    {
        rc = xpu_launch_scan_agg(&kp, P->nregs, (int)rpt, (int)grid, (int)block, lds, nullptr);
        if (rc == 0 && spill)
        {
            const uint32_t sgrid = (uint32_t)g_dev.ncu * (t.blocks_per_cu ? t.blocks_per_cu : 1);
            rc = xpu_launch_spill_agg(&kp, P->nregs, (int)sgrid, (int)block, lds, nullptr);
            if (rc == 0)
                rc = xpu_launch_dump_merge(&kp, P->nregs, (int)kp.dump_nparts, (int)block, lds,
                                           nullptr);
        }
    }
    if (rc != 0)
    {
        hip_ok((hipError_t)rc, eb, "no-alloc kernel launch");
        return ERRCODE_DEVICE_ERROR;
    }
    if (!preagg &&
        xpu_launch_finalize_dst(d_dst, s->d_kds_dst_head, (uint32_t)s->comp.kds_dst_head.size(),
                                dst_length, d_ctl, nullptr) != 0)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no-alloc kernel launch");
        return ERRCODE_DEVICE_ERROR;
    }
    xpu_task_ctl c;
    if (!hip_ok(hipStreamSynchronize(nullptr), eb, "no-alloc pipeline kernel") ||
        !hip_ok(hipMemcpy(&c, d_ctl, sizeof(c), hipMemcpyDeviceToHost), eb, "hipMemcpy"))
        return ERRCODE_DEVICE_ERROR;
    t_last_kernel_ms = 0.0;     /* no events: nothing is timed */
    memset(&t_last_stats, 0, sizeof(t_last_stats));
    t_last_stats.grid = (uint32_t)grid;
    t_last_stats.block = block;
    t_last_stats.ntiles = (uint32_t)ntiles;
    t_last_stats.lds_nents = lds_nents;
    t_last_stats.lds_bytes = lds;
    t_last_stats.rows = nrows;
    t_last_stats.lds_miss = c.lds_miss;
    t_last_stats.global_merges = c.global_merges;
    t_last_stats.spill_rows = c.spill_rows;
    if (res)
    {
        res->nitems_raw = (uint32_t)nrows;
        res->nitems_in = c.nitems_in;
        res->nitems_out = c.nitems_out;
    }
    if (c.errcode && c.errcode != ERRCODE_SUSPEND_NO_SPACE)
    {
        XERR(eb, c.errcode, message_of(s, c.errmsg));
        return (int)c.errcode;
    }
    if (c.errcode == ERRCODE_SUSPEND_NO_SPACE)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, message_of(s, c.errmsg));
        return ERRCODE_DEVICE_ERROR;
    }
    if (!preagg && c.resume_tile != 0)
    {
        uint64_t next = row_begin + (uint64_t)(0xffffffffU - c.resume_tile) * rows_per_tile;
        if (resume_pos)
            *resume_pos = next;
        XERR(eb, ERRCODE_SUSPEND_NO_SPACE, "GpuProjection - no space to write");
        return ERRCODE_SUSPEND_NO_SPACE;
    }
    if (resume_pos)
        *resume_pos = 0;
    return 0;
}

extern "C" int
xpuhip_task_exec(xpuhip_session *s, const kern_data_store *kds_src, size_t src_length,
                 kern_data_store *kds_dst, size_t dst_length, uint64_t *resume_pos,
                 kern_exec_results *res, kern_errorbuf *eb)
{
    int rc;
    if (!s || !kds_src || src_length < sizeof(kern_data_store) || kds_src->length > src_length)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "invalid kds_src");
        return ERRCODE_DEVICE_ERROR;
    }
    const bool tuples = kds_src->format == KDS_FORMAT_BLOCK || kds_src->format == KDS_FORMAT_ROW ||
                        kds_src->format == KDS_FORMAT_HASH;
    if (kds_src->format != KDS_FORMAT_ARROW && !tuples)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "kds_src must be KDS_FORMAT_ARROW, BLOCK, ROW or HASH");
        return ERRCODE_DEVICE_ERROR;
    }
    if (tuples && (src_length < KDS_HEAD_LENGTH(kds_src) || kds_src->nr_colmeta < kds_src->ncols))
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "invalid kds_src");
        return ERRCODE_DEVICE_ERROR;
    }
    if ((rc = ensure_device(eb)) != 0)
        return rc;
    Scratch *x = scratch_get(s, eb);
    if (!x)
        return ERRCODE_DEVICE_ERROR;
    /* group-by and pinned hash results stay on the device */
    bool preagg = s->comp.has_groupby || s->comp.pinned_hash;
    if (!grow(&x->d_src, &x->src_cap, kds_src->length, eb) ||
        (!preagg && !grow(&x->d_dst, &x->dst_cap, dst_length, eb)))
    {
        scratch_put(s, x);
        return ERRCODE_DEVICE_ERROR;
    }
    if (!hip_ok(hipMemcpyAsync(x->d_src, kds_src, kds_src->length, hipMemcpyHostToDevice, x->stream),
                eb, "hipMemcpyAsync(kds_src)"))
    {
        scratch_put(s, x);
        return ERRCODE_DEVICE_ERROR;
    }
    if (tuples)
        rc = task_exec_tuples(s, x, (const kern_data_store *)x->d_src, kds_src,
                              preagg ? nullptr : (kern_data_store *)x->d_dst, dst_length,
                              resume_pos, res, eb);
    else if (s->comp.pinned_hash)
        rc = task_exec_pinned(s, x, (const kern_data_store *)x->d_src, kds_src->nitems, res, eb);
    else
        rc = task_exec_common(s, x, (const kern_data_store *)x->d_src, kds_src->nitems,
                              preagg ? nullptr : (kern_data_store *)x->d_dst, dst_length,
                              resume_pos, res, eb);
    if (!preagg && (rc == 0 || rc == ERRCODE_SUSPEND_NO_SPACE))
    {
        /* header + row index, then the tuples packed at the tail */
        size_t head_len = s->comp.kds_dst_head.size();
        kern_data_store h;
        if (!hip_ok(hipMemcpy(&h, x->d_dst, sizeof(kern_data_store), hipMemcpyDeviceToHost), eb, "hipMemcpy"))
            rc = ERRCODE_DEVICE_ERROR;
        else
        {
            size_t front = head_len + 8ULL * h.nitems;
            if (!hip_ok(hipMemcpy(kds_dst, x->d_dst, front, hipMemcpyDeviceToHost), eb, "hipMemcpy") ||
                (h.usage > 0 &&
                 !hip_ok(hipMemcpy((char *)kds_dst + dst_length - h.usage,
                                   x->d_dst + dst_length - h.usage, h.usage,
                                   hipMemcpyDeviceToHost), eb, "hipMemcpy")))
                rc = ERRCODE_DEVICE_ERROR;
        }
    }
    scratch_put(s, x);
    return rc;
}

/* GPU-Sort of the final buffer's n items: the first nout entries of the
 * sorted row index go to *d_out (in the session's work space) */
static int
sort_final(xpuhip_session *s, uint32_t n, uint32_t nout, const uint64_t *d_ridx,
           const uint64_t **d_out, uint64_t usage, const char **d_tail, uint64_t *used,
           uint32_t *p_nout, kern_errorbuf *eb)
{
    /* with a LIMIT or a window rank the kept tuples are packed
     * (kern_buffer_simple_limit, kern_windowrank_finalize) into a stage
     * sized to their total, which a size pass computes first */
    const xpu_compiled &c = s->comp;
    const bool wr = c.sort_wr_func != 0;
    const bool limited = nout < n || wr;
    (void)usage;
    const uint32_t nw = c.sort_nwords;
    const uint64_t ntiles = (n + XSORT_TILE - 1) / XSORT_TILE;
    const size_t words_bytes = 8ULL * nw * n;
    const size_t need = 2 * words_bytes + 2 * 4ULL * n + 4ULL * 256 * (ntiles + 1) + 16ULL * nw +
                        8ULL * nout + 8 * 256 +     /* 7 slices, each rounded up to 256 B */
                        (limited ? 16ULL * nout + 8 + 3 * 256 : 0) +
                        (wr ? 16ULL * n + 4 * ntiles + 4 + 8ULL * n + 7 * 256 : 0);
    if (need > s->sort_cap)
    {
        if (s->d_sort)
            (void)hipFree(s->d_sort);
        s->d_sort = nullptr;
        s->sort_cap = 0;
        if (!hip_ok(hipMalloc((void **)&s->d_sort, need), eb, "hipMalloc(sort)"))
            return -1;
        s->sort_cap = need;
    }
    char *p = s->d_sort;
    auto take = [&](size_t bytes) { char *r = p; p += (bytes + 255) & ~(size_t)255; return r; };
    uint64_t *W[2] = {(uint64_t *)take(words_bytes), (uint64_t *)take(words_bytes)};
    uint32_t *I[2] = {(uint32_t *)take(4ULL * n), (uint32_t *)take(4ULL * n)};
    uint32_t *counts = (uint32_t *)take(4ULL * 256 * (ntiles + 1));
    uint64_t *wmin = (uint64_t *)take(8ULL * nw);
    uint64_t *wmax = (uint64_t *)take(8ULL * nw);
    uint64_t *ridx_out = (uint64_t *)take(8ULL * nout);
    uint64_t *offs = limited ? (uint64_t *)take(8ULL * nout) : nullptr;
    uint64_t *ridx_packed = limited ? (uint64_t *)take(8ULL * nout) : nullptr;
    uint64_t *d_used = limited ? (uint64_t *)take(8) : nullptr;
    /* window rank: flags, peer counts, keep flags and their prefix, tile
     * sums, the kept count and the kept row index */
    uint32_t *wr_flag = wr ? (uint32_t *)take(4ULL * n) : nullptr;
    uint32_t *wr_cpeer = wr ? (uint32_t *)take(4ULL * n) : nullptr;
    uint32_t *wr_keep = wr ? (uint32_t *)take(4ULL * n) : nullptr;
    uint32_t *wr_pos = wr ? (uint32_t *)take(4ULL * n) : nullptr;
    uint32_t *wr_tiles = wr ? (uint32_t *)take(4ULL * ntiles) : nullptr;
    uint32_t *wr_nkept = wr ? (uint32_t *)take(4) : nullptr;
    uint64_t *wr_ridx = wr ? (uint64_t *)take(8ULL * n) : nullptr;
    if ((size_t)(p - s->d_sort) > s->sort_cap)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "GPU-Sort work space layout exceeds its allocation");
        return -1;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, nullptr);
    xsort_params sp;
    memset(&sp, 0, sizeof(sp));
    sp.kds = s->d_kds_final;
    sp.rowindex = d_ridx;
    sp.n = n;
    sp.nkeys = (uint32_t)c.sort_keys.size();
    sp.nwords = nw;
    for (size_t k = 0; k < c.sort_keys.size(); k++)
    {
        sp.keys[k] = c.sort_keys[k];
        sp.max_anum = c.sort_keys[k].anum > sp.max_anum ? c.sort_keys[k].anum : sp.max_anum;
    }
    sp.words = W[0];
    sp.idx = I[0];
    sp.wmin = wmin;
    sp.wmax = wmax;
    int rc = xpu_launch_fill(wmin, ~0ULL, nw, nullptr);
    if (rc == 0)
        rc = xpu_launch_fill(wmax, 0, nw, nullptr);
    if (rc == 0)
        rc = xpu_launch_sort_keys(&sp, nullptr);
    std::vector<uint64_t> h_min(nw), h_max(nw);
    if (rc != 0 ||
        !hip_ok(hipMemcpy(h_min.data(), wmin, 8ULL * nw, hipMemcpyDeviceToHost), eb, "hipMemcpy") ||
        !hip_ok(hipMemcpy(h_max.data(), wmax, 8ULL * nw, hipMemcpyDeviceToHost), eb, "hipMemcpy"))
    {
        if (rc != 0)
            XERR(eb, ERRCODE_DEVICE_ERROR, "GPU-Sort key pass launch failed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return -1;
    }
    /* least significant digit first: the last word's low byte; a word
     * sorts by the bytes of (greatest - least) only */
    int cur = 0;
    uint32_t passes = 0;
    for (int w = (int)nw - 1; w >= 0 && rc == 0; w--)
    {
        const uint64_t range = h_max[w] - h_min[w];
        for (uint32_t shift = 0; shift < 64 && (range >> shift) != 0 && rc == 0; shift += 8)
        {
            rc = xpu_launch_sort_pass(W[cur], I[cur], W[cur ^ 1], I[cur ^ 1], n, nw, (uint32_t)w,
                                      h_min[w], shift, counts, nullptr);
            cur ^= 1;
            passes++;
        }
    }
    if (rc == 0)
        rc = xpu_launch_sort_permute(d_ridx, I[cur], ridx_out, nout, nullptr);
    /* window rank over the sorted words (W[cur], the last pass's output) */
    const uint64_t *kept_ridx = ridx_out;
    if (rc == 0 && wr)
    {
        rc = xpu_launch_window_rank(W[cur], ridx_out, n, c.sort_wr_pwords, c.sort_wr_owords,
                                    c.sort_wr_func, c.sort_wr_limit, wr_flag, wr_cpeer, wr_keep,
                                    wr_pos, wr_tiles, wr_nkept, wr_ridx, nullptr);
        if (rc == 0 && !hip_ok(hipMemcpy(&nout, wr_nkept, 4, hipMemcpyDeviceToHost), eb, "hipMemcpy"))
        {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return -1;
        }
        kept_ridx = wr_ridx;
    }
    uint64_t stage_len = 0;
    if (rc == 0 && limited)
    {
        rc = xpu_launch_sort_limit_size(s->d_kds_final, kept_ridx, nout, offs, d_used, nullptr);
        if (rc == 0 &&
            (!hip_ok(hipMemcpy(&stage_len, d_used, 8, hipMemcpyDeviceToHost), eb, "hipMemcpy") ||
             !grow(&s->d_sort_stage, &s->sort_stage_cap, stage_len > 0 ? stage_len : 8, eb)))
        {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return -1;
        }
        if (rc == 0)
            rc = xpu_launch_sort_limit_copy(s->d_kds_final, kept_ridx, nout, offs, s->d_sort_stage,
                                            stage_len, ridx_packed, nullptr);
    }
    (void)hipEventRecord(e1, nullptr);
    if (rc != 0 || !hip_ok(hipEventSynchronize(e1), eb, "hipEventSynchronize"))
    {
        if (rc != 0)
            XERR(eb, ERRCODE_DEVICE_ERROR, "GPU-Sort pass launch failed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return -1;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    t_last_stats.sort_ms = ms;
    t_last_stats.sort_items = n;
    t_last_stats.sort_passes = passes;
    *d_out = ridx_out;
    *p_nout = nout;
    if (limited)
    {
        *used = stage_len;
        *d_out = ridx_packed;
        *d_tail = s->d_sort_stage;
    }
    return 0;
}

extern "C" int64_t
xpuhip_task_final(xpuhip_session *s, kern_data_store *out, size_t length,
                  kern_exec_results *res, kern_errorbuf *eb)
{
    if (!s || (!s->comp.has_groupby && !s->comp.pinned_hash))
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "session has no group-by final buffer");
        return -1;
    }
    if (ensure_device(eb) != 0)
        return -1;
    size_t head_len = s->comp.kds_final_head.size();
    std::vector<char> hbuf(head_len);
    if (!hip_ok(hipMemcpy(hbuf.data(), s->d_kds_final, head_len, hipMemcpyDeviceToHost), eb, "hipMemcpy"))
        return -1;
    kern_data_store *h = (kern_data_store *)hbuf.data();
    uint64_t nitems = h->nitems, usage = h->usage, nslots = h->hash_nslots;
    /* GPU-Sort and its LIMIT (gpuservSortingFinalBuffer) */
    const bool sorted = !s->comp.sort_keys.empty();
    uint64_t nout = nitems;
    if (sorted && s->comp.sort_limit > 0 && (uint64_t)s->comp.sort_limit < nitems)
        nout = (uint64_t)s->comp.sort_limit;
    uint64_t need = head_len + 8 * nout + usage;
    if (res)
    {
        memset(res, 0, sizeof(kern_exec_results));
        res->final_plan_task = true;
        res->final_nitems = (uint32_t)nout;
        res->final_usage = usage;
        res->final_total = need;
    }
    if (!out)
        return (int64_t)need;       /* a bound: a LIMIT packs at most this */
    if (length < need)
    {
        XERR(eb, ERRCODE_SUSPEND_NO_SPACE, "kds_final buffer is too small");
        return -1;
    }
    const uint64_t *d_ridx = (const uint64_t *)((char *)s->d_kds_final + head_len + 8 * nslots);
    const char *d_tail = (const char *)s->d_kds_final + s->final_length - usage;
    t_last_stats.sort_ms = 0;
    t_last_stats.sort_items = 0;
    t_last_stats.sort_passes = 0;
    /* a window rank re-packs even one row (kern_windowrank_finalize) */
    if (sorted && (nitems > 1 || (s->comp.sort_wr_func != 0 && nitems > 0)))
    {
        if (nitems > UINT32_MAX)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "GPU-Sort of more than 2^32 - 1 items");
            return -1;
        }
        std::lock_guard<std::mutex> g(s->lock);
        uint32_t kept = (uint32_t)nout;
        if (sort_final(s, (uint32_t)nitems, (uint32_t)nout, d_ridx, &d_ridx, usage, &d_tail,
                       &usage, &kept, eb) != 0)
            return -1;
        nout = kept;                            /* fewer after a window rank */
        need = head_len + 8 * nout + usage;     /* the packed kept tuples */
        if (res)
        {
            res->final_nitems = (uint32_t)nout;
            /* what the written buffer holds (gpu_service.c:5530-5540) */
            res->final_usage = usage;
            res->final_total = need;
            res->final_sorting_msec = (uint32_t)t_last_stats.sort_ms;   /* whole ms, as the reference */
        }
    }
    /* HASH -> ROW on the wire (gpu_service.c:2780-2818) */
    h->format = KDS_FORMAT_ROW;
    h->hash_nslots = 0;
    h->length = need;
    h->nitems = (uint32_t)nout;
    h->usage = usage;
    memcpy(out, hbuf.data(), head_len);
    if (!hip_ok(hipMemcpy((char *)out + head_len, d_ridx,
                          8 * nout, hipMemcpyDeviceToHost), eb, "hipMemcpy") ||
        (usage > 0 &&
         !hip_ok(hipMemcpy((char *)out + need - usage, d_tail, usage, hipMemcpyDeviceToHost), eb,
                 "hipMemcpy")))
        return -1;
    return (int64_t)need;
}

/* The CPU-fallback rows the session's tasks have written since the last
 * call, as the KDS_FORMAT_FALLBACK chunk gpuClientWriteBackPartial sends
 * the backend (gpu_service.c:2853-2889): header, row index, then the
 * kern_fallbackitems at the tail; the device buffer is emptied.  Returns the
 * bytes written, 0 when there are none, with out == NULL the bytes needed,
 * or -1 with eb filled. */
extern "C" int64_t
xpuhip_session_fallback(xpuhip_session *s, kern_data_store *out, size_t length, kern_errorbuf *eb)
{
    if (!s)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no session");
        return -1;
    }
    if (!s->d_kds_fallback)
        return 0;
    if (ensure_device(eb) != 0)
        return -1;
    std::unique_lock<std::shared_mutex> g(s->final_rw);
    const size_t head_len = s->comp.kds_fallback_head.size();
    std::vector<char> hbuf(head_len);
    if (!hip_ok(hipMemcpy(hbuf.data(), s->d_kds_fallback, head_len, hipMemcpyDeviceToHost), eb,
                "hipMemcpy"))
        return -1;
    kern_data_store *h = (kern_data_store *)hbuf.data();
    if (h->nitems == 0)
        return 0;
    uint64_t usage = 0;
    if (!kds_tuple_bytes(s->d_kds_fallback, *h, head_len, &usage, eb, 16))
        return -1;
    const uint64_t dev_len = h->length;
    const uint64_t need = head_len + 8ULL * h->nitems + usage;
    if (!out)
        return (int64_t)need;
    if (length < need)
    {
        XERR(eb, ERRCODE_SUSPEND_NO_SPACE, "fallback buffer of the caller is too small");
        return -1;
    }
    h->length = need;
    h->usage = usage;
    memcpy(out, hbuf.data(), head_len);
    if (!hip_ok(hipMemcpy((char *)out + head_len, (const char *)s->d_kds_fallback + head_len,
                          8ULL * h->nitems, hipMemcpyDeviceToHost), eb, "hipMemcpy") ||
        !hip_ok(hipMemcpy((char *)out + need - usage, (const char *)s->d_kds_fallback + dev_len - usage,
                          usage, hipMemcpyDeviceToHost), eb, "hipMemcpy"))
        return -1;
    /* the device buffer, empty again (gpuClientWriteBackPartial's reset) */
    h->length = dev_len;
    h->nitems = 0;
    h->usage = 0;
    if (!hip_ok(hipMemcpy(s->d_kds_fallback, hbuf.data(), head_len, hipMemcpyHostToDevice), eb,
                "hipMemcpy"))
        return -1;
    return (int64_t)need;
}

extern "C" void
xpuhip_session_close(xpuhip_session *s)
{
    if (!s)
        return;
    for (Scratch *x : s->pool)
        scratch_retire(x);
    s->pool.clear();
    if (s->d_plan) (void)hipFree(s->d_plan);
    if (s->d_const_pool) (void)hipFree(s->d_const_pool);
    if (s->d_kds_dst_head) (void)hipFree(s->d_kds_dst_head);
    if (s->pinned)
    {
        /* the plan's last session frees its pinned buffer */
        std::lock_guard<std::mutex> g(g_pinned_lock);
        if (--s->pinned->refs == 0)
        {
            (void)hipDeviceSynchronize();
            (void)hipFree(s->pinned->d_kds);
            g_pinned.erase(s->pinned->plan_id);
            delete s->pinned;
        }
        s->pinned = nullptr;
        s->d_kds_final = nullptr;
    }
    else if (s->d_kds_final)
        (void)hipFree(s->d_kds_final);
    if (s->d_kds_ovf) (void)hipFree(s->d_kds_ovf);
    if (s->d_kds_fallback) (void)hipFree(s->d_kds_fallback);
    if (s->d_kmrels) (void)hipFree(s->d_kmrels);
    if (s->d_bloom) (void)hipFree(s->d_bloom);
    if (s->d_params) (void)hipFree(s->d_params);
    if (s->d_xids) (void)hipFree(s->d_xids);
    if (s->d_sort) (void)hipFree(s->d_sort);
    if (s->d_sort_stage) (void)hipFree(s->d_sort_stage);
    for (int d = 0; d <= XP_MAX_DEPTH; d++)
        if (s->d_inner_own[d]) (void)hipFree(s->d_inner_own[d]);
    delete s;
}

extern "C" int
xpuhip_session_set_tuning(xpuhip_session *s, const xpuhip_tuning *t)
{
    if (!s || !t)
        return ERRCODE_DEVICE_ERROR;
    xpuhip_tuning d;
    default_tuning(s, &d);
    xpuhip_tuning n = *t;
    if (n.block_size == 0) n.block_size = d.block_size;
    if (n.rows_per_thread == 0) n.rows_per_thread = d.rows_per_thread;
    if (n.blocks_per_cu == 0) n.blocks_per_cu = d.blocks_per_cu;
    if (n.lds_nents == 0) n.lds_nents = d.lds_nents;
    const uint32_t max_block = uses_scan_agg(s) ? XP_AGG_BLOCK_MAX : XP_BLOCK_MAX;
    if (n.block_size % 64 != 0 || n.block_size < 64 || n.block_size > max_block ||
        n.rows_per_thread > 64 || n.blocks_per_cu > 64 || n.jit > XPUHIP_JIT_OFF ||
        n.join_filter > XPUHIP_JOIN_FILTER_OFF || n.join_passes > XPUHIP_JOIN_PASSES_OFF)
        return ERRCODE_DEVICE_ERROR;
    if (uses_scan_agg(s) && n.rows_per_thread != 1 && n.rows_per_thread != 2 &&
        n.rows_per_thread != 4)
        return ERRCODE_DEVICE_ERROR;
    s->tuning = n;
    return 0;
}

extern "C" int
xpuhip_session_get_tuning(xpuhip_session *s, xpuhip_tuning *t)
{
    if (!s || !t)
        return ERRCODE_DEVICE_ERROR;
    *t = s->tuning;
    return 0;
}

extern "C" int
xpuhip_device_alloc(size_t length, void **p, kern_errorbuf *eb)
{
    if (ensure_device(eb) != 0)
        return ERRCODE_DEVICE_ERROR;
    return hip_ok(hipMalloc(p, length), eb, "hipMalloc") ? 0 : ERRCODE_DEVICE_ERROR;
}

extern "C" int
xpuhip_device_free(void *p)
{
    return hipFree(p) == hipSuccess ? 0 : ERRCODE_DEVICE_ERROR;
}

extern "C" int
xpuhip_memcpy_htod(void *d, const void *h, size_t len)
{
    return hipMemcpy(d, h, len, hipMemcpyHostToDevice) == hipSuccess ? 0 : ERRCODE_DEVICE_ERROR;
}

extern "C" int
xpuhip_memcpy_dtoh(void *h, const void *d, size_t len)
{
    return hipMemcpy(h, d, len, hipMemcpyDeviceToHost) == hipSuccess ? 0 : ERRCODE_DEVICE_ERROR;
}

extern "C" int
xpuhip_memcpy_dtod(void *dst, const void *src, size_t len)
{
    kern_errorbuf eb;
    if (ensure_device(&eb) != 0)
        return ERRCODE_DEVICE_ERROR;
    if (xpu_launch_copy(dst, src, len, nullptr) != 0 || hipDeviceSynchronize() != hipSuccess)
        return ERRCODE_DEVICE_ERROR;
    return 0;
}

extern "C" int
xpuhip_device_synchronize(void)
{
    return hipDeviceSynchronize() == hipSuccess ? 0 : ERRCODE_DEVICE_ERROR;
}

extern "C" double
xpuhip_last_kernel_ms(void)
{
    return t_last_kernel_ms;
}

extern "C" int
xpuhip_last_task_stats(xpuhip_task_stats *stats)
{
    if (!stats)
        return ERRCODE_DEVICE_ERROR;
    *stats = t_last_stats;
    return 0;
}

/* ------------------------------------------------------------------
 * session reset and batched launches
 * ------------------------------------------------------------------ */
extern "C" int
xpuhip_session_reset(xpuhip_session *s, kern_errorbuf *eb)
{
    if (!s)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "no session");
        return ERRCODE_DEVICE_ERROR;
    }
    if (!s->comp.has_groupby && !s->comp.pinned_hash)
        return 0;
    std::unique_lock<std::shared_mutex> g(s->final_rw);
    return init_kds_final(s, eb);
}

namespace {

/* device buffers of batched launches, reused across calls (one batch runs
 * at a time in a process); an identical call replays the captured graph */
struct BatchScratch
{
    hipStream_t stream = nullptr;
    hipEvent_t  ev0 = nullptr, ev1 = nullptr;
    xbatch_frag *d_frags = nullptr;     size_t frags_cap = 0;
    xbatch_item *d_items = nullptr;     size_t items_cap = 0;
    xbatch_sitem *d_sitems = nullptr;   size_t sitems_cap = 0;
    xpu_task_ctl *d_ctl = nullptr;      size_t ctl_cap = 0;
    xpu_task_ctl *h_ctl = nullptr;      size_t hctl_cap = 0;
    uint32_t   *d_ticket = nullptr;     size_t ticket_cap = 0;
    uint64_t   *d_spill = nullptr;      size_t spill_cap = 0;
    std::vector<uint64_t> sig;
    hipGraph_t  graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};
std::mutex g_batch_lock;
BatchScratch g_batch;

#define XBATCH_MAX_SITEMS   16384
#define XBATCH_SPILL_BYTES  (8ULL << 30)    /* spill buffer of a batch at most */

void
batch_drop_graph(BatchScratch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->sig.clear();
}

/* grows a device buffer; a reallocation invalidates the captured graph */
bool
batch_grow(BatchScratch *b, char **p, size_t *cap, size_t need, kern_errorbuf *eb)
{
    if (need <= *cap)
        return true;
    batch_drop_graph(b);
    return grow(p, cap, need, eb);
}

bool
same_shape(const xplan *a, const xplan *b)
{
    xplan x, y;
    memcpy(&x, a, sizeof(xplan));
    memcpy(&y, b, sizeof(xplan));
    for (int d = 0; d <= XP_MAX_DEPTH; d++)
        x.depth[d].inner_kds = y.depth[d].inner_kds = 0;
    x.const_pool = y.const_pool = 0;
    if (x.nops != y.nops)
        return false;
    size_t n = offsetof(xplan, ops) + sizeof(xop) * x.nops;
    return memcmp(&x, &y, n) == 0;
}

}   /* namespace */

extern "C" int
xpuhip_batch_exec(const xpuhip_batch_task *tasks, uint32_t ntasks, uint32_t flags,
                  kern_errorbuf *eb)
{
    if (ensure_device(eb) != 0)
        return ERRCODE_DEVICE_ERROR;
    if (!tasks || ntasks == 0)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "empty batch");
        return ERRCODE_DEVICE_ERROR;
    }
    xpuhip_session *s0 = tasks[0].sess;
    for (uint32_t f = 0; f < ntasks; f++)
    {
        xpuhip_session *s = tasks[f].sess;
        if (!s || !uses_scan_agg(s) || !tasks[f].d_kds_src)
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "batch fragments must be scan + group-by sessions with a chunk");
            return ERRCODE_DEVICE_ERROR;
        }
        if (!same_shape(&s->comp.plan, &s0->comp.plan))
        {
            XERR(eb, ERRCODE_DEVICE_ERROR, "batch fragments must share one plan shape");
            return ERRCODE_DEVICE_ERROR;
        }
    }
    const xplan *P = &s0->comp.plan;
    const xpuhip_tuning t = s0->tuning;
    uint32_t block = t.block_size ? t.block_size : 512;
    const uint32_t rpt = t.rows_per_thread >= 4 ? 4 : (t.rows_per_thread == 2 ? 2 : 1);
    const uint32_t grid = (uint32_t)g_dev.ncu * (t.blocks_per_cu ? t.blocks_per_cu : 1);
    const uint32_t lds_nents = P->lds_fast ? t.lds_nents : 0;
    const size_t lds = lds_bytes_for(P, lds_nents);
    if (lds > g_dev.lds_per_block)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "group-by LDS table exceeds the LDS of a block");
        return ERRCODE_DEVICE_ERROR;
    }

    std::lock_guard<std::mutex> g(g_batch_lock);
    BatchScratch *b = &g_batch;
    if (!b->stream &&
        (!hip_ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), eb, "hipStreamCreate") ||
         !hip_ok(hipEventCreate(&b->ev0), eb, "hipEventCreate") ||
         !hip_ok(hipEventCreate(&b->ev1), eb, "hipEventCreate")))
        return ERRCODE_DEVICE_ERROR;

    /* work items: row ranges of about a quarter of a block's fair share
     * (multiples of 4096 rows), in fragment order */
    uint64_t total = 0;
    for (uint32_t f = 0; f < ntasks; f++)
        total += tasks[f].nitems;
    uint64_t item_rows = total / ((uint64_t)grid * 4);
    item_rows = item_rows < (64ULL << 10) ? (64ULL << 10) : (item_rows > (16ULL << 20) ? (16ULL << 20) : item_rows);
    item_rows = (item_rows + 4095) & ~4095ULL;
    std::vector<xbatch_item> items;
    for (uint32_t f = 0; f < ntasks; f++)
        for (uint64_t r = 0; r < tasks[f].nitems; r += item_rows)
        {
            xbatch_item it;
            it.frag = f;
            it.row_begin = (uint32_t)r;
            it.row_end = (uint32_t)(r + item_rows < tasks[f].nitems ? r + item_rows : tasks[f].nitems);
            it.__pad = 0;
            items.push_back(it);
        }

    /* spill regions: an eighth of each fragment's rows (at least 64K records),
     * scaled down to the batch's spill budget */
    const bool spill = P->spill_ok && lds_nents > 0;
    const uint64_t rec = 8ULL * (1 + P->spill_nregs);
    std::vector<uint64_t> caps(ntasks, 0);
    uint64_t spill_recs = 0;
    if (spill)
    {
        for (uint32_t f = 0; f < ntasks; f++)
        {
            caps[f] = tasks[f].nitems / 8 > 65536 ? tasks[f].nitems / 8 : 65536;
            spill_recs += caps[f];
        }
        if (spill_recs * rec > XBATCH_SPILL_BYTES)
        {
            double k = (double)XBATCH_SPILL_BYTES / (double)(spill_recs * rec);
            spill_recs = 0;
            for (uint32_t f = 0; f < ntasks; f++)
            {
                caps[f] = (uint64_t)(caps[f] * k);
                spill_recs += caps[f];
            }
        }
    }

    std::vector<uint64_t> sig;
    sig.push_back(ntasks);
    sig.push_back(flags);
    sig.push_back(((uint64_t)block << 32) | rpt);
    sig.push_back(((uint64_t)grid << 32) | lds_nents);
    sig.push_back(t.jit);
    for (uint32_t f = 0; f < ntasks; f++)
    {
        sig.push_back((uint64_t)(uintptr_t)tasks[f].sess);
        sig.push_back((uint64_t)(uintptr_t)tasks[f].d_kds_src);
        sig.push_back(tasks[f].nitems);
    }

    const JitFns *jf = nullptr;
    if (t.jit != XPUHIP_JIT_OFF && (t.jit == XPUHIP_JIT_ON || total >= XPUHIP_JIT_AUTO_ROWS))
    {
        jf = jit_function(s0, XJ_SCAN_AGG, (int)rpt, eb);
        if (!jf && t.jit == XPUHIP_JIT_ON)
            return ERRCODE_DEVICE_ERROR;
    }
    /* the interpreter's kernels take blocks of up to XP_BLOCK_MAX threads */
    if (!jf && block > XP_BLOCK_MAX)
        block = XP_BLOCK_MAX;

    const bool replay = (flags & XPUHIP_BATCH_GRAPH) && b->gexec && b->sig == sig;
    if (!replay)
    {
        if (!batch_grow(b, (char **)&b->d_frags, &b->frags_cap, sizeof(xbatch_frag) * ntasks, eb) ||
            !batch_grow(b, (char **)&b->d_items, &b->items_cap, sizeof(xbatch_item) * items.size(), eb) ||
            !batch_grow(b, (char **)&b->d_sitems, &b->sitems_cap, sizeof(xbatch_sitem) * XBATCH_MAX_SITEMS, eb) ||
            !batch_grow(b, (char **)&b->d_ctl, &b->ctl_cap, sizeof(xpu_task_ctl) * ntasks, eb) ||
            !batch_grow(b, (char **)&b->d_ticket, &b->ticket_cap, 64, eb) ||
            (spill && !batch_grow(b, (char **)&b->d_spill, &b->spill_cap, spill_recs * rec, eb)))
            return ERRCODE_DEVICE_ERROR;
        if (b->hctl_cap < sizeof(xpu_task_ctl) * ntasks)
        {
            if (b->h_ctl)
                (void)hipHostFree(b->h_ctl);
            b->h_ctl = nullptr;
            b->hctl_cap = 0;
            if (!hip_ok(hipHostMalloc((void **)&b->h_ctl, sizeof(xpu_task_ctl) * ntasks,
                                      hipHostMallocDefault), eb, "hipHostMalloc"))
                return ERRCODE_DEVICE_ERROR;
            b->hctl_cap = sizeof(xpu_task_ctl) * ntasks;
        }
        std::vector<xbatch_frag> frags(ntasks);
        uint64_t off = 0;
        for (uint32_t f = 0; f < ntasks; f++)
        {
            xpuhip_session *s = tasks[f].sess;
            xbatch_frag &fr = frags[f];
            memset(&fr, 0, sizeof(fr));
            fr.kds_src = tasks[f].d_kds_src;
            fr.kds_final = s->d_kds_final;
            fr.ctl = b->d_ctl + f;
            fr.const_pool = s->d_const_pool;
            fr.params = s->d_params;
            fr.param_nulls = s->comp.param_nulls;
            fr.spill = spill ? b->d_spill + off * (rec / 8) : nullptr;
            fr.spill_cap = caps[f];
            fr.nrows = tasks[f].nitems;
            off += caps[f];
        }
        if (!hip_ok(hipMemcpy(b->d_frags, frags.data(), sizeof(xbatch_frag) * ntasks,
                              hipMemcpyHostToDevice), eb, "hipMemcpy") ||
            !hip_ok(hipMemcpy(b->d_items, items.data(), sizeof(xbatch_item) * items.size(),
                              hipMemcpyHostToDevice), eb, "hipMemcpy"))
            return ERRCODE_DEVICE_ERROR;
        batch_drop_graph(b);
    }

    xpu_kparams kp;
    memset(&kp, 0, sizeof(kp));
    kp.plan = s0->d_plan;
    kp.const_pool = s0->d_const_pool;
    kp.lds_nents = lds_nents;
    kp.batch_frags = b->d_frags;
    kp.batch_items = b->d_items;
    kp.batch_sitems = b->d_sitems;
    kp.batch_ticket = b->d_ticket;
    kp.batch_nfrags = ntasks;
    kp.batch_nitems = (uint32_t)items.size();
    kp.batch_nsitems_cap = XBATCH_MAX_SITEMS;

    /* the launch sequence: clear control blocks and tickets, pass 1, spill
     * setup, pass 2 (no memset nodes: capturable) */
    auto enqueue = [&]() -> int {
        int rc = xpu_launch_fill(b->d_ctl, 0, sizeof(xpu_task_ctl) * ntasks / 8, b->stream);
        if (!rc)
            rc = xpu_launch_fill(b->d_ticket, 0, 8, b->stream);
        if (rc)
            return rc;
        if (jf)
        {
            void *args[] = {&kp};
            rc = (int)hipModuleLaunchKernel(jf->batch, grid, 1, 1, block, 1, 1, (unsigned)lds,
                                            b->stream, args, nullptr);
            if (!rc)
                rc = xpu_launch_batch_setup(&kp, b->stream);
            if (!rc)
                rc = (int)hipModuleLaunchKernel(jf->batch_spill, grid, 1, 1,
                                                block < XP_BLOCK_MAX ? block : XP_BLOCK_MAX, 1, 1,
                                                (unsigned)lds, b->stream, args, nullptr);
            return rc;
        }
        return xpu_launch_batch(&kp, P->nregs, (int)rpt, grid, block, lds, b->stream);
    };

    int rc = 0;
    if (flags & XPUHIP_BATCH_GRAPH)
    {
        if (!replay)
        {
            if (!hip_ok(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), eb,
                        "hipStreamBeginCapture"))
                return ERRCODE_DEVICE_ERROR;
            rc = enqueue();
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &graph);
            if (rc || e != hipSuccess)
            {
                if (graph)
                    (void)hipGraphDestroy(graph);
                hip_ok(rc ? (hipError_t)rc : e, eb, "batch graph capture");
                return ERRCODE_DEVICE_ERROR;
            }
            if (!hip_ok(hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0), eb,
                        "hipGraphInstantiate"))
            {
                (void)hipGraphDestroy(graph);
                return ERRCODE_DEVICE_ERROR;
            }
            b->graph = graph;
            b->sig = sig;
        }
        (void)hipEventRecord(b->ev0, b->stream);
        rc = (int)hipGraphLaunch(b->gexec, b->stream);
        (void)hipEventRecord(b->ev1, b->stream);
    }
    else
    {
        (void)hipEventRecord(b->ev0, b->stream);
        rc = enqueue();
        (void)hipEventRecord(b->ev1, b->stream);
    }
    if (rc)
    {
        hip_ok((hipError_t)rc, eb, "batch launch");
        return ERRCODE_DEVICE_ERROR;
    }
    (void)hipMemcpyAsync(b->h_ctl, b->d_ctl, sizeof(xpu_task_ctl) * ntasks, hipMemcpyDeviceToHost,
                         b->stream);
    if (!hip_ok(hipStreamSynchronize(b->stream), eb, "batch"))
        return ERRCODE_DEVICE_ERROR;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    t_last_kernel_ms = ms;
    memset(&t_last_stats, 0, sizeof(t_last_stats));
    t_last_stats.kernel_ms = ms;
    t_last_stats.pass1_ms = ms;
    t_last_stats.grid = grid;
    t_last_stats.block = block;
    t_last_stats.ntiles = (uint32_t)items.size();
    t_last_stats.lds_nents = lds_nents;
    t_last_stats.lds_bytes = lds;
    t_last_stats.rows = total;
    t_last_stats.jit = jf ? 1 : 0;
    t_last_stats.join_slices = 0;
    t_last_stats.heap_ms = 0.0;
    t_last_stats.heap_rows = 0;

    int first_err = 0;
    for (uint32_t f = 0; f < ntasks; f++)
    {
        const xpu_task_ctl *c = b->h_ctl + f;
        t_last_stats.lds_miss += c->lds_miss;
        t_last_stats.global_merges += c->global_merges;
        t_last_stats.spill_rows += c->spill_rows;
        if (tasks[f].results)
        {
            kern_exec_results *r = tasks[f].results;
            memset(r, 0, sizeof(*r));
            r->nitems_raw = tasks[f].nitems;
            r->nitems_in = c->nitems_in;
            r->nitems_out = c->nitems_out;
        }
        if (c->errcode && !first_err)
        {
            char msg[200];
            snprintf(msg, sizeof(msg), "batch fragment %u: %s", f,
                     message_of(tasks[f].sess, c->errmsg));
            XERR(eb, ERRCODE_DEVICE_ERROR, msg);
            first_err = ERRCODE_DEVICE_ERROR;
        }
    }
    return first_err;
}

/* ------------------------------------------------------------------
 * autotuner: launch shapes tried on the caller's own data
 * ------------------------------------------------------------------ */
namespace {

/* tuned shapes by plan shape (plan bytes without session addresses), so
 * that later sessions of the same shape start tuned */
std::mutex g_tune_lock;
std::map<std::string, xpuhip_tuning> g_tuned;

std::string
shape_key(const xplan *P)
{
    xplan x;
    memcpy(&x, P, sizeof(xplan));
    for (int d = 0; d <= XP_MAX_DEPTH; d++)
        x.depth[d].inner_kds = 0;
    x.const_pool = 0;
    return std::string((const char *)&x, offsetof(xplan, ops) + sizeof(xop) * x.nops);
}

/* candidate launch shapes for a session's task kind */
std::vector<xpuhip_tuning>
tuning_candidates(const xpuhip_session *s)
{
    const xplan *P = &s->comp.plan;
    std::vector<xpuhip_tuning> c;
    auto add = [&](uint32_t block, uint32_t rpt, uint32_t bpc, double lds_frac) {
        xpuhip_tuning t;
        memset(&t, 0, sizeof(t));
        t.block_size = block;
        t.rows_per_thread = rpt;
        t.blocks_per_cu = bpc;
        t.jit = s->tuning.jit;
        size_t budget = g_dev.lds_per_block > 4096 ? g_dev.lds_per_block - 4096 : 0;
        t.lds_nents = lds_nents_within(P, (size_t)(budget * lds_frac / bpc));
        if (P->lds_fast && t.lds_nents == 0)
            return;
        c.push_back(t);
    };
    if (uses_scan_agg(s))
    {
        add(1024, 2, 1, 1.0);
        add(1024, 4, 1, 1.0);
        add(512, 4, 1, 1.0);
        add(512, 2, 1, 1.0);
        add(256, 4, 1, 1.0);
        add(512, 4, 1, 0.75);
        add(512, 4, 2, 1.0);
    }
    else if (s->comp.has_groupby)
    {
        add(512, 4, 2, 0.3);
        add(256, 4, 2, 0.3);
        add(512, 4, 1, 0.3);
        add(256, 4, 4, 0.3);
        add(512, 2, 2, 0.3);
    }
    else
    {
        add(256, 1, 2, 0.0);
        add(256, 1, 4, 0.0);
        add(512, 1, 2, 0.0);
        add(512, 1, 4, 0.0);
    }
    return c;
}

}   /* namespace */

extern "C" int
xpuhip_session_autotune(xpuhip_session *s, const kern_data_store *d_src, uint32_t nitems,
                        kern_data_store *d_dst, size_t dst_length, xpuhip_tuning *best,
                        kern_errorbuf *eb)
{
    if (!s || !d_src)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "autotune needs a session and a chunk");
        return ERRCODE_DEVICE_ERROR;
    }
    std::vector<xpuhip_tuning> cands = tuning_candidates(s);
    xpuhip_tuning keep = s->tuning, win = s->tuning;
    double win_ms = 1e30;
    for (const xpuhip_tuning &t : cands)
    {
        if (xpuhip_session_set_tuning(s, &t) != 0)
            continue;
        double ms = 1e30;
        for (int rep = 0; rep < 2; rep++)      /* the first run compiles / warms */
        {
            uint64_t pos = 0;
            int rc = xpuhip_task_exec_device(s, d_src, nitems, d_dst, dst_length, &pos, nullptr, eb);
            if (rc == ERRCODE_SUSPEND_NO_SPACE)
                rc = 0;         /* a projection that overflows still measures */
            if (rc != 0)
            {
                s->tuning = keep;
                xpuhip_session_reset(s, eb);
                return rc;
            }
            ms = t_last_kernel_ms;
        }
        if (ms < win_ms)
        {
            win_ms = ms;
            win = s->tuning;
        }
    }
    s->tuning = win;
    {
        std::lock_guard<std::mutex> g(g_tune_lock);
        g_tuned[shape_key(&s->comp.plan)] = win;
    }
    if (best)
        *best = win;
    return xpuhip_session_reset(s, eb);
}

extern "C" int
xpuhip_batch_autotune(const xpuhip_batch_task *tasks, uint32_t ntasks, xpuhip_tuning *best,
                      kern_errorbuf *eb)
{
    if (!tasks || ntasks == 0 || !tasks[0].sess)
    {
        XERR(eb, ERRCODE_DEVICE_ERROR, "empty batch");
        return ERRCODE_DEVICE_ERROR;
    }
    xpuhip_session *s0 = tasks[0].sess;
    std::vector<xpuhip_tuning> cands = tuning_candidates(s0);
    xpuhip_tuning win = s0->tuning;
    double win_ms = 1e30;
    for (const xpuhip_tuning &t : cands)
    {
        if (xpuhip_session_set_tuning(s0, &t) != 0)
            continue;
        double ms = 1e30;
        for (int rep = 0; rep < 2; rep++)
        {
            int rc = xpuhip_batch_exec(tasks, ntasks, 0, eb);
            if (rc != 0)
                return rc;
            ms = t_last_kernel_ms;
        }
        if (ms < win_ms)
        {
            win_ms = ms;
            win = s0->tuning;
        }
    }
    {
        std::lock_guard<std::mutex> g(g_tune_lock);
        g_tuned[shape_key(&s0->comp.plan)] = win;
    }
    for (uint32_t f = 0; f < ntasks; f++)
    {
        tasks[f].sess->tuning = win;
        int rc = xpuhip_session_reset(tasks[f].sess, eb);
        if (rc)
            return rc;
    }
    if (best)
        *best = win;
    return 0;
}

/* the tuned shape of an earlier session of the same plan, if any */
bool
xpuhip_tuned_shape(const xplan *P, xpuhip_tuning *t)
{
    std::lock_guard<std::mutex> g(g_tune_lock);
    auto it = g_tuned.find(shape_key(P));
    if (it == g_tuned.end())
        return false;
    *t = it->second;
    return true;
}
