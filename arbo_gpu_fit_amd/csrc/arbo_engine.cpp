// arbo_engine.cpp — the C ABI of include/arbo_c.h over the HIP kernels.
//
// The engine owns all device memory of a batch and reuses it for the whole
// run (the reference allocates and frees ~50 buffers per MCMC step,
// main.cpp:198-249 and 284-335).  A day is two kernel launches plus a
// one-thread day-counter increment, all on the engine's stream; with
// ARBO_F_GRAPH the day loop is captured once into a hipGraph of several days
// and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "arbo_c.h"
#include "arbo_device.h"
#include "arbo_tables.h"

using namespace arbo;

namespace {

constexpr uint32_t kGraphDays = 8;

// A device allocation that a later batch reuses when it is large enough, so
// an MCMC driver re-initialising the engine every step does not reallocate.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;     // elements in use by the current batch
    size_t cap = 0;   // elements allocated
};

struct SimHost {
    uint32_t C = 0, S = 0;
    uint32_t comm_off = 0;
    uint64_t n_humans = 0, n_slots = 0;
    std::vector<uint32_t> humans, slots;
    std::vector<uint64_t> h_off, m_off;    // global padded offsets per community
    uint64_t series_off = 0;
    uint64_t spatial_off = 0;
};

}  // namespace

struct ArboEngine {
    int device = 0;
    uint32_t flags = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool ready = false;
    uint32_t day = 0, max_days = 0, n_sims = 0, n_comm = 0;
    std::vector<SimHost> sims;
    std::vector<SimDesc> sim_desc;
    // device memory
    DevBuf<CommDesc> comm;
    DevBuf<SimDesc> simd;
    DevBuf<uint32_t> h_word, alive, cnt, nld, active, hsum, series, spatial, csr_thr, csr_inbox, in_edge, in_slot, tables, day_ctr;
    DevBuf<uint16_t> h_next, m_word, m_next;
    DevBuf<int32_t> h_birth, hist;
    DevBuf<uint8_t> aoi_mark;
    DevBuf<uint32_t> aoi, aoi_pos;
    DevBuf<unsigned long long> lam, scanned;
    DevBuf<uint32_t> part;   // per-community partials of the day
    DevBuf<float> dens;
    uint64_t series_total = 0;
    EngineView view{};
    // graph of kGraphDays days
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // timing
    std::vector<hipEvent_t> ev_pool, ev_used;
    double kernel_ms[2] = {0.0, 0.0};
    Tuning tuning{4, 2, 0, 8, 0, 1};   // the autotuner's usual pick on large batches
    uint64_t kernel_launches[2] = {0, 0};
};

namespace {

int fail(ArboEngine* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    return code;
}

#define HIP_TRY(e, call)                                                                   \
    do {                                                                                   \
        hipError_t _st = (call);                                                           \
        if (_st != hipSuccess)                                                             \
            return fail((e), ARBO_E_DEVICE, std::string(#call ": ") + hipGetErrorString(_st)); \
    } while (0)

template <class T>
void dev_free(DevBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    b.cap = 0;
}

template <class T>
int dev_alloc(ArboEngine* e, DevBuf<T>& b, size_t n) {
    if (b.p && n <= b.cap) {
        b.n = n;
        return ARBO_OK;
    }
    dev_free(b);
    if (n == 0) return ARBO_OK;
    hipError_t st = hipMalloc((void**)&b.p, n * sizeof(T));
    if (st != hipSuccess) {
        b.p = nullptr;
        return fail(e, ARBO_E_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(st));
    }
    b.n = n;
    b.cap = n;
    return ARBO_OK;
}

void release_graph(ArboEngine* e) {
    if (e->graph_exec) { (void)hipGraphExecDestroy(e->graph_exec); e->graph_exec = nullptr; }
    if (e->graph) { (void)hipGraphDestroy(e->graph); e->graph = nullptr; }
}

void free_batch(ArboEngine* e) {
    release_graph(e);
    dev_free(e->comm); dev_free(e->simd);
    dev_free(e->h_word); dev_free(e->alive); dev_free(e->cnt); dev_free(e->nld); dev_free(e->active); dev_free(e->hsum); dev_free(e->series); dev_free(e->spatial); dev_free(e->csr_thr); dev_free(e->csr_inbox);
    dev_free(e->in_edge); dev_free(e->in_slot); dev_free(e->tables); dev_free(e->day_ctr);
    dev_free(e->h_next); dev_free(e->m_word); dev_free(e->m_next); dev_free(e->h_birth); dev_free(e->hist); dev_free(e->aoi_mark); dev_free(e->aoi); dev_free(e->aoi_pos);
    dev_free(e->lam); dev_free(e->scanned); dev_free(e->dens); dev_free(e->part);
    e->ready = false;
}

uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Append a table to a pool, reusing an identical earlier copy.
template <class T>
uint32_t pool_add(std::vector<T>& pool, std::map<std::vector<T>, uint32_t>& seen, const std::vector<T>& t) {
    auto it = seen.find(t);
    if (it != seen.end()) return it->second;
    const uint32_t off = (uint32_t)pool.size();
    pool.insert(pool.end(), t.begin(), t.end());
    seen.emplace(t, off);
    return off;
}

int record_timing_event(ArboEngine* e, hipEvent_t* out) {
    if (e->ev_pool.empty()) {
        hipEvent_t ev;
        HIP_TRY(e, hipEventCreate(&ev));
        e->ev_pool.push_back(ev);
    }
    *out = e->ev_pool.back();
    e->ev_pool.pop_back();
    e->ev_used.push_back(*out);
    HIP_TRY(e, hipEventRecord(*out, e->stream));
    return ARBO_OK;
}

int drain_timing(ArboEngine* e) {
    if (e->ev_used.empty()) return ARBO_OK;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    for (size_t i = 0; i + 2 < e->ev_used.size(); i += 3) {
        float a = 0.0f, b = 0.0f;
        HIP_TRY(e, hipEventElapsedTime(&a, e->ev_used[i], e->ev_used[i + 1]));
        HIP_TRY(e, hipEventElapsedTime(&b, e->ev_used[i + 1], e->ev_used[i + 2]));
        e->kernel_ms[0] += a;
        e->kernel_ms[1] += b;
        e->kernel_launches[0]++;
        e->kernel_launches[1]++;
    }
    for (hipEvent_t ev : e->ev_used) e->ev_pool.push_back(ev);
    e->ev_used.clear();
    return ARBO_OK;
}

// The kernels of one day, at day_offset days past the device counter.
int launch_day_kernels(ArboEngine* e, uint32_t day_offset) {
    EngineView v = e->view;
    v.day_offset = day_offset;
    HIP_TRY(e, launch_demography(v, e->tuning, e->stream));
    HIP_TRY(e, launch_longdist(v, e->tuning, e->stream));
    HIP_TRY(e, launch_events(v, e->tuning, e->stream));
    return ARBO_OK;
}

int launch_day(ArboEngine* e) {
    if (int rc = launch_day_kernels(e, 0)) return rc;
    HIP_TRY(e, launch_advance(e->day_ctr.p, 1, e->stream));
    return ARBO_OK;
}

int launch_day_timed(ArboEngine* e) {
    hipEvent_t a, b, c;
    int rc = record_timing_event(e, &a);
    if (rc) return rc;
    HIP_TRY(e, launch_demography(e->view, e->tuning, e->stream));
    if ((rc = record_timing_event(e, &b))) return rc;
    HIP_TRY(e, launch_longdist(e->view, e->tuning, e->stream));
    HIP_TRY(e, launch_events(e->view, e->tuning, e->stream));
    if ((rc = record_timing_event(e, &c))) return rc;
    HIP_TRY(e, launch_advance(e->day_ctr.p, 1, e->stream));
    // bound the number of outstanding events
    if (e->ev_used.size() >= 3 * 512) return drain_timing(e);
    return ARBO_OK;
}

int build_graph(ArboEngine* e) {
    HIP_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    // the block's days read the counter at its start plus their offset, and
    // one launch advances it past the block
    int rc = ARBO_OK;
    for (uint32_t d = 0; d < kGraphDays && rc == ARBO_OK; ++d) rc = launch_day_kernels(e, d);
    if (rc == ARBO_OK && launch_advance(e->day_ctr.p, kGraphDays, e->stream) != hipSuccess)
        rc = fail(e, ARBO_E_DEVICE, "launch_advance in capture");
    hipGraph_t g = nullptr;
    hipError_t st = hipStreamEndCapture(e->stream, &g);
    if (rc) return rc;
    if (st != hipSuccess) return fail(e, ARBO_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(st));
    e->graph = g;
    HIP_TRY(e, hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0));
    return ARBO_OK;
}

// Copy the members of every community of one simulation between the padded
// device layout and a dense host array (H = T to read, const T to write).
template <class T, class H>
int copy_members(ArboEngine* e, const SimHost& s, T* dev, const std::vector<uint64_t>& off,
                 const std::vector<uint32_t>& count, H* host, bool to_device) {
    if (s.C == 0) return ARBO_OK;
    const uint64_t lo = off[0];
    const uint64_t hi = off[s.C - 1] + count[s.C - 1];
    std::vector<T> buf(hi - lo);
    HIP_TRY(e, hipMemcpy(buf.data(), dev + lo, buf.size() * sizeof(T), hipMemcpyDeviceToHost));
    uint64_t k = 0;
    for (uint32_t c = 0; c < s.C; ++c) {
        T* seg = buf.data() + (off[c] - lo);
        if (to_device) std::memcpy(seg, host + k, count[c] * sizeof(T));
        else std::memcpy((void*)(host + k), seg, count[c] * sizeof(T));
        k += count[c];
    }
    if (to_device) HIP_TRY(e, hipMemcpy(dev + lo, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice));
    return ARBO_OK;
}


}  // namespace

extern "C" {

int arbo_create(const ArboEngineDesc* desc, ArboEngine** out) {
    if (!out) return ARBO_E_INVALID;
    *out = nullptr;
    if (!desc || desc->abi_version != ARBO_ABI_VERSION) return ARBO_E_INVALID;
    ArboEngine* e = new ArboEngine();
    e->device = desc->device;
    e->flags = desc->flags;
    hipError_t st = hipSetDevice(e->device);
    if (st == hipSuccess) st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (st != hipSuccess) {
        delete e;
        return ARBO_E_DEVICE;
    }
    *out = e;
    return ARBO_OK;
}

int arbo_batch_init(ArboEngine* e, uint32_t n_sims, const ArboShape* shapes, const ArboParams* params,
                    const uint64_t* seeds, uint32_t max_days) {
    if (!e) return ARBO_E_INVALID;
    if (n_sims == 0 || !shapes || !params || !seeds) return fail(e, ARBO_E_INVALID, "empty batch");
    if (max_days == 0 || max_days > (1u << 23)) return fail(e, ARBO_E_INVALID, "max_days must be 1..2^23");
    for (uint32_t i = 0; i < n_sims; ++i) {
        std::string why = validate(shapes[i], params[i], max_days);
        if (!why.empty()) return fail(e, ARBO_E_INVALID, "simulation " + std::to_string(i) + ": " + why);
    }
    HIP_TRY(e, hipSetDevice(e->device));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    release_graph(e);   // the graph captured the old buffers' addresses
    e->ready = false;

    // host tables, built in parallel over simulations
    std::vector<SimTables> tabs(n_sims);
    {
        const uint32_t nt = std::max(1u, std::min<uint32_t>(n_sims, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (uint32_t w = 0; w < nt; ++w)
            pool.emplace_back([&, w]() {
                for (uint32_t i = w; i < n_sims; i += nt) build_tables(shapes[i], params[i], seeds[i], max_days, tabs[i]);
            });
        for (auto& th : pool) th.join();
    }

    // layout
    e->n_sims = n_sims;
    e->max_days = max_days;
    e->sims.assign(n_sims, SimHost());
    e->sim_desc.assign(n_sims, SimDesc());
    std::vector<CommDesc> comm;
    std::vector<uint32_t> csr_thr, csr_inbox, in_edge, in_slot, table_pool;
    std::vector<float> dens_pool;
    std::map<std::vector<uint32_t>, uint32_t> seen_u32;
    std::map<std::vector<float>, uint32_t> seen_f32;
    uint64_t h_tot = 0, m_tot = 0, ser_tot = 0, spa_tot = 0;
    uint32_t max_n = 0, max_cap = 0, death_m_lds = 1;
    for (uint32_t i = 0; i < n_sims; ++i) {
        const SimTables& t = tabs[i];
        const ArboParams& p = params[i];
        SimHost& sh = e->sims[i];
        sh.C = t.C;
        sh.S = t.S;
        sh.comm_off = (uint32_t)comm.size();
        sh.n_humans = t.n_humans;
        sh.n_slots = t.n_slots;
        sh.humans = t.humans;
        sh.slots = t.slots;
        sh.h_off.resize(t.C);
        sh.m_off.resize(t.C);
        sh.series_off = ser_tot;
        ser_tot += series_words(max_days, t.S);
        sh.spatial_off = spa_tot;
        if (e->flags & ARBO_F_SPATIAL) spa_tot += ((uint64_t)max_days + 1) * t.S * t.C;

        SimDesc& sd = e->sim_desc[i];
        std::memset(&sd, 0, sizeof sd);
        sd.k0 = (uint32_t)seeds[i];
        sd.k1 = (uint32_t)(seeds[i] >> 32);
        sd.C = t.C;
        sd.S = t.S;
        sd.comm_off = sh.comm_off;
        sd.beta_hm = p.beta_hm;
        sd.beta_mh = p.beta_mh;
        sd.omega = p.long_distance;
        sd.recovery = p.recovery;
        sd.eip = p.eip;
        sd.intro = p.external_infection;
        sd.iip_min = p.iip_min;
        sd.iip_span = p.iip_max - p.iip_min + 1;
        sd.death_h_off = pool_add(table_pool, seen_u32, t.death_h);
        sd.death_m_off = pool_add(table_pool, seen_u32, t.death_m);
        sd.init_h_off = pool_add(table_pool, seen_u32, t.init_h);
        sd.init_m_off = pool_add(table_pool, seen_u32, t.init_m);
        sd.immune_off = pool_add(table_pool, seen_u32, t.immune);
        sd.dens_off = pool_add(dens_pool, seen_f32, t.dens);
        sd.climate = t.climate ? 1u : 0u;
        if (t.climate) sd.clim_off = pool_add(dens_pool, seen_f32, t.clim);
        sd.seed_h = t.seed_h;
        sd.seed_m = t.seed_m;
        sd.death_m_max = t.death_m_max;
        sd.death_m_guide_off = pool_add(table_pool, seen_u32, t.death_m_guide);
        sd.count_law = p.count_law;
        sd.no_distance = p.no_distance;
        death_m_lds = std::max(death_m_lds, t.death_m_max + 1);
        sd.series_off = sh.series_off;
        sd.spatial_off = sh.spatial_off;

        const uint32_t row_base = (uint32_t)csr_thr.size();
        for (uint32_t k = 0; k < t.thr.size(); ++k) csr_thr.push_back(t.thr[k] | (t.alias[k] << 27));
        // inbox row of every CSR entry: (global destination) * 9 + (index of
        // the source among the destination's in-edges)
        csr_inbox.resize(csr_thr.size());
        for (uint32_t d = 0; d < t.C; ++d)
            for (uint32_t e = t.in_ptr[d]; e < t.in_ptr[d + 1]; ++e)
                csr_inbox[row_base + t.row_ptr[t.in_src[e]] + t.in_slot[e]] =
                    (sh.comm_off + d) * kInboxEdges + (e - t.in_ptr[d]);
        const uint32_t in_base = (uint32_t)in_edge.size();
        for (uint32_t k = 0; k < t.in_src.size(); ++k) {
            in_edge.push_back(t.in_src[k] + sh.comm_off);
            in_slot.push_back(t.in_slot[k]);
        }
        uint32_t hb = 0, mb = 0;
        for (uint32_t c = 0; c < t.C; ++c) {
            CommDesc cd;
            std::memset(&cd, 0, sizeof cd);
            h_tot = align_up(h_tot, kHumanAlign);
            m_tot = align_up(m_tot, kMosqAlign);
            cd.h_off = h_tot;
            cd.m_off = m_tot;
            sh.h_off[c] = h_tot;
            sh.m_off[c] = m_tot;
            cd.sim = i;
            cd.c = c;
            cd.n = t.humans[c];
            cd.cap = t.slots[c];
            cd.h_base = hb;
            cd.m_base = mb;
            cd.row_off = row_base + t.row_ptr[c];
            cd.row_len = t.row_ptr[c + 1] - t.row_ptr[c];
            cd.in_off = in_base + t.in_ptr[c];
            cd.in_len = t.in_ptr[c + 1] - t.in_ptr[c];
            cd.base = t.base[c];
            cd.flags = params[i].external_infection > 0.0f ? kCommIntro : 0u;
            if (params[i].no_distance > 0.0f) {   // the column of the row that is the community itself
                uint32_t self = 0;
                while (self < cd.row_len && t.to[t.row_ptr[c] + self] != c) ++self;
                if (self == cd.row_len) return fail(e, ARBO_E_INVALID, "community missing from its own CSR row");
                cd.flags |= kCommStay | (self << kCommSelfShift);
            }
            comm.push_back(cd);
            h_tot += cd.n;
            m_tot += cd.cap;
            hb += cd.n;
            mb += cd.cap;
            max_n = std::max(max_n, cd.n);
            max_cap = std::max(max_cap, cd.cap);
        }
        e->sim_desc[i].h_begin = sh.h_off[0];
        e->sim_desc[i].h_end = align_up(h_tot, kHumanAlign);
    }
    h_tot = align_up(h_tot, kHumanAlign) + kHumanAlign;
    m_tot = align_up(m_tot, kMosqAlign) + ARBO_MAX_COMMUNITY_SLOTS;  // room for whole-workgroup vector reads
    if (comm.size() >= (1ull << 31)) return fail(e, ARBO_E_INVALID, "too many communities in the batch");
    e->n_comm = (uint32_t)comm.size();
    e->series_total = ser_tot;

    int rc;
    if ((rc = dev_alloc(e, e->comm, comm.size())) || (rc = dev_alloc(e, e->simd, n_sims)) ||
        (rc = dev_alloc(e, e->h_word, h_tot)) || (rc = dev_alloc(e, e->h_next, h_tot)) ||
        (rc = dev_alloc(e, e->h_birth, h_tot)) ||
        (rc = dev_alloc(e, e->hist, (e->flags & (ARBO_F_HISTORY | ARBO_F_AOI)) ? 4 * h_tot : 0)) ||
        (rc = dev_alloc(e, e->aoi_mark, (e->flags & ARBO_F_AOI) ? h_tot : 0)) ||
        (rc = dev_alloc(e, e->aoi, (e->flags & ARBO_F_AOI) ? (size_t)n_sims * 4 * kAoiRecord : 0)) ||
        (rc = dev_alloc(e, e->aoi_pos, (e->flags & ARBO_F_AOI) ? (size_t)n_sims * 4 : 0)) || (rc = dev_alloc(e, e->m_word, m_tot)) ||
        (rc = dev_alloc(e, e->m_next, m_tot)) || (rc = dev_alloc(e, e->alive, m_tot / 32)) || (rc = dev_alloc(e, e->cnt, comm.size() * kInboxEdges * 8)) || (rc = dev_alloc(e, e->nld, comm.size() * 8)) ||
        (rc = dev_alloc(e, e->hsum, comm.size() * 4)) || (rc = dev_alloc(e, e->active, comm.size())) ||
        (rc = dev_alloc(e, e->lam, (size_t)2 * n_sims * 8)) || (rc = dev_alloc(e, e->scanned, (size_t)2 * n_sims)) || (rc = dev_alloc(e, e->part, comm.size() * kPartWords)) || (rc = dev_alloc(e, e->series, ser_tot)) || (rc = dev_alloc(e, e->spatial, spa_tot)) ||
        (rc = dev_alloc(e, e->csr_thr, csr_thr.size())) || (rc = dev_alloc(e, e->csr_inbox, csr_inbox.size())) || (rc = dev_alloc(e, e->in_edge, in_edge.size())) ||
        (rc = dev_alloc(e, e->in_slot, in_slot.size())) || (rc = dev_alloc(e, e->tables, table_pool.size())) ||
        (rc = dev_alloc(e, e->dens, dens_pool.size())) || (rc = dev_alloc(e, e->day_ctr, 1))) {
        free_batch(e);
        return rc;
    }
    HIP_TRY(e, hipMemcpy(e->comm.p, comm.data(), comm.size() * sizeof(CommDesc), hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->simd.p, e->sim_desc.data(), n_sims * sizeof(SimDesc), hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->csr_thr.p, csr_thr.data(), csr_thr.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->csr_inbox.p, csr_inbox.data(), csr_inbox.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->in_edge.p, in_edge.data(), in_edge.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->in_slot.p, in_slot.data(), in_slot.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->tables.p, table_pool.data(), table_pool.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(e, hipMemcpy(e->dens.p, dens_pool.data(), dens_pool.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(e, launch_fill(e->series.p, ser_tot, 0u, e->stream));
    if (spa_tot) HIP_TRY(e, launch_fill(e->spatial.p, spa_tot, 0u, e->stream));
    HIP_TRY(e, launch_fill(reinterpret_cast<uint32_t*>(e->lam.p), (uint64_t)2 * n_sims * 8 * 2, 0u, e->stream));
    HIP_TRY(e, launch_fill(reinterpret_cast<uint32_t*>(e->scanned.p), (uint64_t)2 * n_sims * 2, 0u, e->stream));
    HIP_TRY(e, launch_fill(e->day_ctr.p, 1, 0u, e->stream));
    HIP_TRY(e, launch_fill(e->nld.p, comm.size() * 8, 0u, e->stream));
    HIP_TRY(e, launch_fill(e->hsum.p, comm.size() * 4, 0u, e->stream));   // every community scans its humans on day 0
    HIP_TRY(e, launch_fill(e->active.p, comm.size(), 0u, e->stream));
    // padding words: humans beyond a community are never read as members, but
    // keep the whole array defined
    HIP_TRY(e, launch_fill(e->h_word.p, h_tot, 0u, e->stream));
    if (e->flags & (ARBO_F_HISTORY | ARBO_F_AOI))
        HIP_TRY(e, launch_fill(reinterpret_cast<uint32_t*>(e->hist.p), 4 * h_tot, 0xFFFFFFFFu, e->stream));
    if (e->flags & ARBO_F_AOI) {   // h_tot is a multiple of 4: the marks fill whole words
        HIP_TRY(e, launch_fill(reinterpret_cast<uint32_t*>(e->aoi_mark.p), h_tot / 4, 0u, e->stream));
        HIP_TRY(e, launch_fill(e->aoi.p, (uint64_t)n_sims * 4 * kAoiRecord, 0u, e->stream));
        HIP_TRY(e, launch_fill(e->aoi_pos.p, (uint64_t)n_sims * 4, 0u, e->stream));
    }
    HIP_TRY(e, launch_fill(reinterpret_cast<uint32_t*>(e->m_word.p), m_tot / 2, 0x00030003u, e->stream));

    EngineView& v = e->view;
    v.comm = e->comm.p;
    v.sims = e->simd.p;
    v.h_word = e->h_word.p;
    v.h_next = e->h_next.p;
    v.h_birth = e->h_birth.p;
    v.m_word = e->m_word.p;
    v.m_next = e->m_next.p;
    v.alive = e->alive.p;
    v.hsum = e->hsum.p;
    v.active = e->active.p;
    v.cnt = e->cnt.p;
    v.nld = e->nld.p;
    v.lam = e->lam.p;
    v.part = e->part.p;
    v.scanned = e->scanned.p;
    v.series = e->series.p;
    v.spatial = (e->flags & ARBO_F_SPATIAL) ? e->spatial.p : nullptr;
    v.hist = (e->flags & (ARBO_F_HISTORY | ARBO_F_AOI)) ? e->hist.p : nullptr;
    v.aoi_mark = (e->flags & ARBO_F_AOI) ? e->aoi_mark.p : nullptr;
    v.aoi = e->aoi.p;
    v.aoi_pos = e->aoi_pos.p;
    v.aoi_from = max_days > kAoiWindow ? max_days - kAoiWindow : 0u;
    v.csr_inbox = e->csr_inbox.p;
    v.csr_thr = e->csr_thr.p;
    v.in_edge = e->in_edge.p;
    v.in_slot = e->in_slot.p;
    v.tables = e->tables.p;
    v.dens = e->dens.p;
    v.day = e->day_ctr.p;
    v.n_sims = n_sims;
    v.n_comm = e->n_comm;
    v.max_days = max_days;
    v.death_m_lds = death_m_lds;
    v.max_words = (max_cap + 31) >> 5;   // above 256 the day kernels take their eight-wave forms
    {   // about one community per thread of k_day_reduce, at most kMaxReduceSplit workgroups per
        // simulation and about 2,048 in all (8 per CU)
        uint32_t max_c = 1;
        for (const SimHost& sh : e->sims) max_c = std::max(max_c, sh.C);
        const uint32_t cap = std::max<uint32_t>(1, std::min<uint32_t>(kMaxReduceSplit, 2048 / n_sims));
        v.reduce_split = std::min<uint32_t>(cap, (max_c + 255) / 256);
    }

    HIP_TRY(e, launch_init(v, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    e->day = 0;
    e->ready = true;
    if (e->flags & ARBO_F_GRAPH) {
        if ((rc = build_graph(e))) return rc;
    }
    return ARBO_OK;
}

int arbo_batch_step(ArboEngine* e, uint32_t n_days) {
    if (!e) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    if ((uint64_t)e->day + n_days > e->max_days) return fail(e, ARBO_E_CAPACITY, "series capacity exceeded");
    HIP_TRY(e, hipSetDevice(e->device));
    uint32_t done = 0;
    int rc;
    if ((e->flags & ARBO_F_TIMING) == 0 && e->graph_exec) {
        while (n_days - done >= kGraphDays) {
            HIP_TRY(e, hipGraphLaunch(e->graph_exec, e->stream));
            done += kGraphDays;
        }
    }
    for (; done < n_days; ++done) {
        rc = (e->flags & ARBO_F_TIMING) ? launch_day_timed(e) : launch_day(e);
        if (rc) return rc;
    }
    e->day += n_days;
    return ARBO_OK;
}

namespace {

Tuning from_abi(const ArboTuning& t) {
    return Tuning{t.census_unroll, t.early_mosq, t.xcd_order, t.ld_split, t.waves_per_simd, t.waves_per_comm};
}
ArboTuning to_abi(const Tuning& t) { return ArboTuning{t.unroll, t.early_m, t.xcd, t.ld_split, t.wps, t.nw}; }

// Sets the launch knobs.  The captured graph holds the old kernels: it is
// released, and rebuilt at once unless `rebuild` is false (the autotuner
// launches its trial days directly and rebuilds once, for the tuning it
// keeps).
int apply_tuning(ArboEngine* e, const Tuning& tu, bool rebuild = true) {
    if (!tuning_supported(tu)) return fail(e, ARBO_E_INVALID, "unsupported tuning");
    HIP_TRY(e, hipSetDevice(e->device));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    e->tuning = tu;
    release_graph(e);
    if (rebuild && e->ready && (e->flags & ARBO_F_GRAPH)) return build_graph(e);
    return ARBO_OK;
}

}  // namespace

int arbo_tuning_candidates(ArboTuning* out, uint32_t cap, uint32_t* n) {
    if (!n) return ARBO_E_INVALID;
    uint32_t k = 0;
    for (const auto& d : kDemographyShapes)
        for (uint32_t x = 0; x <= 3; x += 3) {
            if (out && k < cap) out[k] = ArboTuning{d[0], d[1], x, 8, 0, 1};
            ++k;
        }
    for (const auto& d : kOccShapes) {   // eight waves per SIMD: for batches of few communities
        if (out && k < cap) out[k] = ArboTuning{d[0], d[1], 0, 8, 8, 1};
        ++k;
    }
    for (const auto& d : kWideShapes) {   // several waves per community: for batches of few communities
        if (out && k < cap) out[k] = ArboTuning{d[0], d[1], 0, 8, 0, d[2]};
        ++k;
    }
    *n = k;
    return ARBO_OK;
}

int arbo_set_tuning(ArboEngine* e, const ArboTuning* t) {
    if (!e || !t) return ARBO_E_INVALID;
    return apply_tuning(e, from_abi(*t));
}

int arbo_get_tuning(const ArboEngine* e, ArboTuning* out) {
    if (!e || !out) return ARBO_E_INVALID;
    *out = to_abi(e->tuning);
    return ARBO_OK;
}

int arbo_autotune(ArboEngine* e, uint32_t days_each, ArboTuning* chosen, double* ms_per_day) {
    if (!e || days_each == 0) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    uint32_t n = 0;
    arbo_tuning_candidates(nullptr, 0, &n);
    std::vector<ArboTuning> cand(n);
    arbo_tuning_candidates(cand.data(), n, &n);
    if ((uint64_t)e->day + (uint64_t)n * days_each > e->max_days)
        return fail(e, ARBO_E_CAPACITY, "autotuning needs candidates x days_each days of series capacity");
    HIP_TRY(e, hipSetDevice(e->device));
    // The candidates take turns one day at a time, in forward order on even
    // rounds and backward on odd ones, so that every candidate's days have the
    // same mean and a steady drift of the epidemic favours none.  Events
    // bracket the demography kernel and the rest of the day; the demography
    // knobs come from the candidate with the least demography time, the event
    // kernel's order bit from the one with the least rest-of-day time.
    const uint32_t ne = 3 * n;
    std::vector<hipEvent_t> ev(ne, nullptr);
    int rc = ARBO_OK;
    for (uint32_t i = 0; i < ne && rc == ARBO_OK; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) rc = fail(e, ARBO_E_DEVICE, "autotune: hipEventCreate");
    std::vector<double> ta(n, 0.0), tb(n, 0.0);
    for (uint32_t r = 0; r < days_each && rc == ARBO_OK; ++r) {
        hipError_t st = hipSuccess;
        for (uint32_t j = 0; j < n && rc == ARBO_OK && st == hipSuccess; ++j) {
            const uint32_t k = (r & 1u) ? n - 1 - j : j;
            if ((rc = apply_tuning(e, from_abi(cand[k]), false))) break;
            st = hipEventRecord(ev[3 * k], e->stream);
            if (st == hipSuccess) st = launch_demography(e->view, e->tuning, e->stream);
            if (st == hipSuccess) st = hipEventRecord(ev[3 * k + 1], e->stream);
            if (st == hipSuccess) st = launch_longdist(e->view, e->tuning, e->stream);
            if (st == hipSuccess) st = launch_events(e->view, e->tuning, e->stream);
            if (st == hipSuccess) st = hipEventRecord(ev[3 * k + 2], e->stream);
            if (st == hipSuccess) st = launch_advance(e->day_ctr.p, 1, e->stream);
            if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
            float x = 0.0f, y = 0.0f;
            if (st == hipSuccess) st = hipEventElapsedTime(&x, ev[3 * k], ev[3 * k + 1]);
            if (st == hipSuccess) st = hipEventElapsedTime(&y, ev[3 * k + 1], ev[3 * k + 2]);
            if (st == hipSuccess) {
                ta[k] += x;
                tb[k] += y;
                e->day += 1;
            }
        }
        if (rc == ARBO_OK && st != hipSuccess) rc = fail(e, ARBO_E_DEVICE, std::string("autotune: ") + hipGetErrorString(st));
    }
    uint32_t ka = 0, kb = 0;
    for (uint32_t k = 1; k < n; ++k) {
        if (ta[k] < ta[ka]) ka = k;
        if (tb[k] < tb[kb]) kb = k;
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (rc) {   // leave a usable engine: the tuning before the failure, with its graph
        const int rc2 = apply_tuning(e, e->tuning);
        (void)rc2;
        return rc;
    }
    Tuning keep = from_abi(cand[ka]);
    keep.xcd = (keep.xcd & 1u) | (cand[kb].xcd_order & 2u);
    if ((rc = apply_tuning(e, keep))) return rc;
    if (chosen) *chosen = to_abi(keep);
    if (ms_per_day) *ms_per_day = (ta[ka] + tb[kb]) / days_each;
    return ARBO_OK;
}

int arbo_sync(ArboEngine* e) {
    if (!e) return ARBO_E_INVALID;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return ARBO_OK;
}

int arbo_days_done(const ArboEngine* e, uint32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    *out = e->day;
    return ARBO_OK;
}

int arbo_sim_sizes(const ArboEngine* e, uint32_t sim, uint32_t* communities, uint64_t* humans, uint64_t* slots,
                   uint32_t* serotypes) {
    if (!e || !e->ready || sim >= e->n_sims) return ARBO_E_INVALID;
    const SimHost& s = e->sims[sim];
    if (communities) *communities = s.C;
    if (humans) *humans = s.n_humans;
    if (slots) *slots = s.n_slots;
    if (serotypes) *serotypes = s.S;
    return ARBO_OK;
}

int arbo_read_series(ArboEngine* e, uint32_t sim, ArboSeries* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->n_sims) return fail(e, ARBO_E_INVALID, "bad simulation index");
    const SimHost& s = e->sims[sim];
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    std::vector<uint32_t> blk(series_words(e->max_days, s.S));
    HIP_TRY(e, hipMemcpy(blk.data(), e->series.p + s.series_off, blk.size() * 4, hipMemcpyDeviceToHost));
    const SeriesIdx si = series_index(0, e->max_days, s.S);
    const size_t rows = (size_t)e->day + 1;
    if (out->n_count) std::memcpy(out->n_count, &blk[si.n_count], rows * 4);
    if (out->m_count) std::memcpy(out->m_count, &blk[si.m_count], rows * 4);
    if (out->n_infected) std::memcpy(out->n_infected, &blk[si.n_inf], rows * s.S * 4);
    if (out->n_incidence) std::memcpy(out->n_incidence, &blk[si.n_inc], rows * s.S * 4);
    if (out->n_one_subpop) std::memcpy(out->n_one_subpop, &blk[si.n_one], rows * s.S * 4);
    return ARBO_OK;
}


int arbo_read_spatial(ArboEngine* e, uint32_t sim, uint32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->n_sims) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!(e->flags & ARBO_F_SPATIAL)) return fail(e, ARBO_E_STATE, "engine created without ARBO_F_SPATIAL");
    const SimHost& s = e->sims[sim];
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    const size_t words = ((size_t)e->day + 1) * s.S * s.C;
    HIP_TRY(e, hipMemcpy(out, e->spatial.p + s.spatial_off, words * 4, hipMemcpyDeviceToHost));
    return ARBO_OK;
}

int arbo_read_history(ArboEngine* e, uint32_t sim, int32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->n_sims) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!(e->flags & (ARBO_F_HISTORY | ARBO_F_AOI)))
        return fail(e, ARBO_E_STATE, "engine created without ARBO_F_HISTORY");
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    const SimHost& s = e->sims[sim];
    if (s.C == 0) return ARBO_OK;
    // the simulation's padded block of 4 words per human, then [s][i] in community order
    const uint64_t lo = s.h_off[0], hi = s.h_off[s.C - 1] + s.humans[s.C - 1];
    std::vector<int32_t> blk(4 * (hi - lo));
    HIP_TRY(e, hipMemcpy(blk.data(), e->hist.p + 4 * lo, blk.size() * 4, hipMemcpyDeviceToHost));
    const uint64_t N = s.n_humans;
    uint64_t k = 0;
    for (uint32_t c = 0; c < s.C; ++c)
        for (uint32_t i = 0; i < s.humans[c]; ++i, ++k)
            for (uint32_t v = 0; v < s.S; ++v) out[v * N + k] = blk[4 * (s.h_off[c] - lo + i) + v];
    return ARBO_OK;
}

int arbo_climate_day(const ArboClimate* c, uint32_t row, double* out) {
    if (!c || !out || !c->temperature || !c->humidity || !c->precipitation || row >= c->days) return ARBO_E_INVALID;
    climate_day(*c, row, out);
    return ARBO_OK;
}

int arbo_read_aoi(ArboEngine* e, uint32_t sim, uint32_t* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->n_sims) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!(e->flags & ARBO_F_AOI)) return fail(e, ARBO_E_STATE, "engine created without ARBO_F_AOI");
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    HIP_TRY(e, hipMemcpy(out, e->aoi.p + (size_t)sim * 4 * kAoiRecord, 4 * kAoiRecord * 4, hipMemcpyDeviceToHost));
    return ARBO_OK;
}

int arbo_read_state(ArboEngine* e, uint32_t sim, ArboState* out) {
    if (!e || !out) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->n_sims) return fail(e, ARBO_E_INVALID, "bad simulation index");
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    const SimHost& s = e->sims[sim];
    int rc;
    if (out->h_word && (rc = copy_members(e, s, e->h_word.p, s.h_off, s.humans, out->h_word, false))) return rc;
    if (out->h_next && (rc = copy_members(e, s, e->h_next.p, s.h_off, s.humans, out->h_next, false))) return rc;
    if (out->h_birth && (rc = copy_members(e, s, e->h_birth.p, s.h_off, s.humans, out->h_birth, false))) return rc;
    if (out->m_word && (rc = copy_members(e, s, e->m_word.p, s.m_off, s.slots, out->m_word, false))) return rc;
    if (out->m_next && (rc = copy_members(e, s, e->m_next.p, s.m_off, s.slots, out->m_next, false))) return rc;
    return ARBO_OK;
}

int arbo_write_state(ArboEngine* e, uint32_t sim, const ArboState* in) {
    if (!e || !in) return ARBO_E_INVALID;
    if (!e->ready || sim >= e->n_sims) return fail(e, ARBO_E_INVALID, "bad simulation index");
    if (!in->h_word || !in->h_next || !in->h_birth || !in->m_word || !in->m_next)
        return fail(e, ARBO_E_INVALID, "every census array is required");
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    const SimHost& s = e->sims[sim];
    int rc;
    if ((rc = copy_members(e, s, e->h_word.p, s.h_off, s.humans, in->h_word, true))) return rc;
    if ((rc = copy_members(e, s, e->h_next.p, s.h_off, s.humans, in->h_next, true))) return rc;
    if ((rc = copy_members(e, s, e->h_birth.p, s.h_off, s.humans, in->h_birth, true))) return rc;
    if ((rc = copy_members(e, s, e->m_word.p, s.m_off, s.slots, in->m_word, true))) return rc;
    if ((rc = copy_members(e, s, e->m_next.p, s.m_off, s.slots, in->m_next, true))) return rc;
    // the restored humans are scanned on the next day
    HIP_TRY(e, launch_fill(e->hsum.p, e->hsum.n, 0u, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return ARBO_OK;
}

int arbo_set_day(ArboEngine* e, uint32_t day) {
    if (!e) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    if (day > e->max_days) return fail(e, ARBO_E_CAPACITY, "day beyond series capacity");
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    HIP_TRY(e, hipMemcpy(e->day_ctr.p, &day, 4, hipMemcpyHostToDevice));
    std::vector<uint32_t> blk(e->series_total);
    HIP_TRY(e, hipMemcpy(blk.data(), e->series.p, blk.size() * 4, hipMemcpyDeviceToHost));
    for (const SimHost& s : e->sims) {
        const SeriesIdx si = series_index(s.series_off, e->max_days, s.S);
        for (uint64_t r = (uint64_t)day + 1; r <= e->max_days; ++r) {
            blk[si.n_count + r] = 0;
            blk[si.m_count + r] = 0;
            for (uint32_t v = 0; v < s.S; ++v) {
                blk[si.n_inf + r * s.S + v] = 0;
                blk[si.n_inc + r * s.S + v] = 0;
                blk[si.n_one + r * s.S + v] = 0;
            }
        }
    }
    HIP_TRY(e, hipMemcpy(e->series.p, blk.data(), blk.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(e, launch_fill(reinterpret_cast<uint32_t*>(e->lam.p), (uint64_t)e->lam.n * 2, 0u, e->stream));
    HIP_TRY(e, launch_fill(e->hsum.p, e->hsum.n, 0u, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    e->day = day;
    return ARBO_OK;
}

int arbo_kernel_time(ArboEngine* e, uint32_t kind, double* ms, uint64_t* launches) {
    if (!e || kind > 1) return ARBO_E_INVALID;
    int rc = drain_timing(e);
    if (rc) return rc;
    if (ms) *ms = e->kernel_ms[kind];
    if (launches) *launches = e->kernel_launches[kind];
    return ARBO_OK;
}

int arbo_reset_timing(ArboEngine* e) {
    if (!e) return ARBO_E_INVALID;
    int rc = drain_timing(e);
    if (rc) return rc;
    e->kernel_ms[0] = e->kernel_ms[1] = 0.0;
    e->kernel_launches[0] = e->kernel_launches[1] = 0;
    if (e->ready) {
        HIP_TRY(e, hipSetDevice(e->device));
        HIP_TRY(e, launch_fill(reinterpret_cast<uint32_t*>(e->scanned.p), (uint64_t)e->scanned.n * 2, 0u, e->stream));
    }
    return ARBO_OK;
}

int arbo_scan_counts(ArboEngine* e, uint64_t* human_vectors, uint64_t* mosquito_vectors) {
    if (!e) return ARBO_E_INVALID;
    if (!e->ready) return fail(e, ARBO_E_STATE, "arbo_batch_init has not been called");
    HIP_TRY(e, hipSetDevice(e->device));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    std::vector<unsigned long long> v(e->scanned.n);
    HIP_TRY(e, hipMemcpy(v.data(), e->scanned.p, v.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    uint64_t h = 0, m = 0;
    for (size_t i = 0; i + 1 < v.size(); i += 2) {
        h += v[i];
        m += v[i + 1];
    }
    if (human_vectors) *human_vectors = h;
    if (mosquito_vectors) *mosquito_vectors = m;
    return ARBO_OK;
}

void arbo_destroy(ArboEngine* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    free_batch(e);
    for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : e->ev_used) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* arbo_last_error(const ArboEngine* e) { return e ? e->err.c_str() : "null engine"; }

}  // extern "C"
