// Serving loop: one resident cohort, solve requests admitted into the columns of one
// shared block PCG as slots and HBM free up (continuous batching).
//
// Memory: one record pool sized by the HBM budget holds every admitted request's
// record-space data -- its record structure (individual and time per record, individual
// blocks, LDS tiles) and its six PCG vectors (b, x, r, z, p, V p) -- as one contiguous
// range of nrec + 1 elements rounded up to 64, placed first-fit. Column j of the block
// reads its range through a per-column offset (RecMeta::col_off), and every record-space
// kernel loops over that column's own records, so a request costs the HBM and the
// record-space work of its own size. A request is admitted (FIFO) when a slot is free
// and a pool range of its size is.
//
// Time: admissions are staged one step ahead. After the step's iteration is enqueued,
// the next requests' data is uploaded on a copy stream while the GPU iterates, and the
// step returns without waiting for the iteration. The next step waits for it, takes the
// columns it finished out of the block, joins the staged requests (the compute stream
// waits on the uploads' event, then starts each new column's recurrence) and enqueues
// its own iteration; only then does it copy out the finished columns' solutions (on the
// copy stream, while the GPU iterates) and return their ranges to the pool.
//
// The genotype contraction takes all occupied columns at once; its fixed-point scale,
// the PCG scalars and the record structure are per column, and the reductions of a column
// are split over a grid fixed by the per-request record cap, so a request's arithmetic
// does not depend on which other requests share the block.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>

#include "context.h"

struct ServeState {
    struct Request {
        int64_t id = 0, nrec = 0;
        int32_t nblk = 0, ntiles = 0;
        std::vector<int32_t> indiv, rec_blk, blk_ind, tile_blk;
        std::vector<int64_t> blk_off;
        std::vector<double> times, y;
        double theta[5] = {0, 0, 0, 0, 0};
        double t_submit = 0;
        int64_t span = 0;   // pool elements
    };
    struct Done {
        std::vector<double> x;
        globe_serve_result res{};
    };
    // kRetiring: finished in the last iteration; its column holds no records any more and
    // its x is copied out while the next iteration runs, then the slot is free
    enum SlotState { kFree = 0, kStaged = 1, kBusy = 2, kRetiring = 3 };
    struct Retire {
        int slot = 0;
        globe_serve_result res{};
    };
    globe_serve_opts o{};
    int slots = 0;
    int64_t cap = 0;        // records of the largest request
    int64_t pool = 0;       // pool elements
    globe::DevBuf B, X, R, Z, P, VP;                          // [pool] doubles
    globe::DevBuf indiv, times, rec_blk, blk_off, blk_ind, tile_blk;  // [pool]
    globe::DevBuf col_off, ntiles_col, nrec_col, theta, state;
    globe::DevBuf part_a, part_b, part_c;
    hipStream_t copy = nullptr;   // uploads of staged requests
    hipEvent_t staged_ev = nullptr;
    std::map<int64_t, int64_t> free_ranges;  // pool offset -> length
    // the queue and the id counter are shared with globe_serve_submit, which may run on
    // another thread (request intake) while the loop steps
    std::mutex qmu;
    std::deque<Request> queue;
    std::vector<Request> slot_req;
    std::vector<int> slot_state;
    std::vector<int64_t> slot_off;
    std::vector<int> admit_iter;
    std::vector<double> t_admit;
    std::map<int64_t, Done> done;
    int64_t next_id = 0;
    int global_iter = 0;
    int64_t peak_in_use = 0, waits_for_room = 0;
    bool staged_pending = false;
    bool inflight = false;  // an iteration enqueued by the last step, not yet waited for
    // rows of the genotype product the block needs: the busy requests' individuals in
    // ascending order, padded to a multiple of 512 with individuals outside them (whose
    // rows of G S0 no record reads); rows_n = 0: every row
    std::vector<int32_t> rows_host;
    globe::DevBuf rows_dev;   // n + 512 entries, allocated at open
    int64_t rows_n = 0;
    bool rows_dirty = true, rows_on = true;
    // per individual: how many staged, busy or retiring requests hold it; union_cnt: how
    // many individuals some request holds (a request's counts are added when it is
    // staged and removed when its slot frees, so the rows the gather takes are a superset
    // of the busy requests' individuals, which changes no row's value)
    bool count_rows = false;
    std::vector<int32_t> ind_refs;
    int64_t union_cnt = 0;
    std::chrono::steady_clock::time_point t0;
    globe::PcgState* h_state = nullptr;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    // record structure of columns j0 .. (pooled)
    globe::RecMeta meta(int j0) const {
        globe::RecMeta m;
        m.indiv = indiv.as<int32_t>();
        m.times = times.as<double>();
        m.rec_blk = rec_blk.as<int32_t>();
        m.blk_off = blk_off.as<int64_t>();
        m.blk_ind = blk_ind.as<int32_t>();
        m.tile_blk = tile_blk.as<int32_t>();
        m.col_off = col_off.as<int64_t>() + j0;
        m.ntiles_col = ntiles_col.as<int32_t>() + j0;
        // the launch grid covers the most tiles a column of cap records can have
        m.ntiles = (int32_t)globe::max_tiles(cap);
        m.nrec_col = nrec_col.as<int64_t>() + j0;
        m.nrec = cap;
        m.cap = globe::kTileRecords;
        return m;
    }
    // first-fit range of `len` pool elements, or -1
    int64_t alloc(int64_t len) {
        for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
            if (it->second < len) continue;
            const int64_t off = it->first, rest = it->second - len;
            free_ranges.erase(it);
            if (rest > 0) free_ranges.emplace(off + len, rest);
            return off;
        }
        return -1;
    }
    void add_rows(const Request& r, int d) {
        if (!count_rows) return;
        if (d > 0) {
            for (const int32_t i : r.blk_ind)
                if (ind_refs[(size_t)i]++ == 0) ++union_cnt;
        } else {
            for (const int32_t i : r.blk_ind)
                if (--ind_refs[(size_t)i] == 0) --union_cnt;
        }
        rows_dirty = true;
    }
    void release(int64_t off, int64_t len) {
        auto it = free_ranges.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_ranges.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_ranges.erase(nx);
        }
        if (it != free_ranges.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_ranges.erase(it);
            }
        }
    }
};

namespace globe {

namespace {

// device bytes per pool element: six FP64 vectors and the record structure (individual,
// time, block of each record; block offsets, block individuals, tile starts)
constexpr int64_t kPoolElemBytes = 6 * 8 + 4 + 8 + 4 + 8 + 4 + 4;

// pool elements a request of nrec records occupies: its records, one more block offset,
// rounded to 64 so that every range starts 512-byte aligned
int64_t request_span(int64_t nrec) { return round_up(nrec + 1, 64); }

ServeState& server(globe_ctx* c) {
    require(c->serve != nullptr, GLOBE_E_STATE, "serving loop not open (globe_serve_open)");
    return *c->serve;
}

void prepare(globe_ctx* c, const ServeState& S, const globe_request* q, ServeState::Request& r) {
    require(q != nullptr && q->nrec >= 1, GLOBE_E_INVALID, "empty request");
    require(q->nrec <= S.cap, GLOBE_E_INVALID,
            "request has " + std::to_string(q->nrec) + " records; a slot holds " + std::to_string(S.cap));
    require(q->indiv && q->times && q->y, GLOBE_E_INVALID, "null request arrays");
    require(q->theta[0] > 0.0, GLOBE_E_INVALID, "sig2e must be positive");
    const int64_t nrec = q->nrec;
    r.nrec = nrec;
    r.indiv.assign(q->indiv, q->indiv + nrec);
    r.times.assign(q->times, q->times + nrec);
    r.y.assign(q->y, q->y + nrec);
    std::memcpy(r.theta, q->theta, sizeof(r.theta));
    r.rec_blk.resize(nrec);
    std::vector<char> seen((size_t)c->n, 0);
    int64_t in_tile = 0, b = -1;
    r.tile_blk.assign(1, 0);
    for (int64_t i = 0; i < nrec; ++i) {
        const int32_t ind = r.indiv[i];
        require(ind >= 0 && ind < c->n, GLOBE_E_INVALID, "request individual outside the cohort");
        if (i == 0 || r.indiv[i - 1] != ind) {
            require(!seen[ind], GLOBE_E_INVALID, "records of an individual must be contiguous");
            seen[ind] = 1;
            if (b >= 0) {
                const int64_t len = i - r.blk_off.back();
                require(len <= kTileMaxRecords, GLOBE_E_INVALID, "individual with too many records");
            }
            ++b;
            r.blk_off.push_back(i);
            r.blk_ind.push_back(ind);
        }
        r.rec_blk[i] = (int32_t)b;
    }
    r.blk_off.push_back(nrec);
    r.nblk = (int32_t)r.blk_ind.size();
    require(nrec - r.blk_off[r.nblk - 1] <= kTileMaxRecords, GLOBE_E_INVALID,
            "individual with too many records");
    for (int32_t q2 = 0; q2 < r.nblk; ++q2) {
        const int64_t len = r.blk_off[q2 + 1] - r.blk_off[q2];
        if (in_tile > 0 && (in_tile + len > kTileRecords || q2 - r.tile_blk.back() == kTileBlocks)) {
            r.tile_blk.push_back(q2);
            in_tile = 0;
        }
        in_tile += len;
    }
    r.tile_blk.push_back(r.nblk);
    r.ntiles = (int32_t)r.tile_blk.size() - 1;
    r.span = request_span(nrec);
}

// Reserve a slot and a pool range for the queue's head and upload its data on the copy
// stream (it joins the block at the next admit_staged).
bool stage_next(ServeState& S) {
    int j = 0;
    while (j < S.slots && S.slot_state[j] != ServeState::kFree) ++j;
    if (j == S.slots) return false;
    ServeState::Request r;
    int64_t off = -1;
    {
        std::lock_guard<std::mutex> lk(S.qmu);
        if (S.queue.empty()) return false;
        off = S.alloc(S.queue.front().span);
        if (off < 0) {  // first in, first admitted: the head waits for room
            ++S.waits_for_room;
            return false;
        }
        r = std::move(S.queue.front());
        S.queue.pop_front();
    }
    const int64_t nrec = r.nrec;
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S.copy));
    };
    h2d(S.indiv.as<int32_t>() + off, r.indiv.data(), sizeof(int32_t) * nrec);
    h2d(S.times.as<double>() + off, r.times.data(), sizeof(double) * nrec);
    h2d(S.rec_blk.as<int32_t>() + off, r.rec_blk.data(), sizeof(int32_t) * nrec);
    h2d(S.blk_off.as<int64_t>() + off, r.blk_off.data(), sizeof(int64_t) * r.blk_off.size());
    h2d(S.blk_ind.as<int32_t>() + off, r.blk_ind.data(), sizeof(int32_t) * r.blk_ind.size());
    h2d(S.tile_blk.as<int32_t>() + off, r.tile_blk.data(), sizeof(int32_t) * r.tile_blk.size());
    h2d(S.B.as<double>() + off, r.y.data(), sizeof(double) * nrec);
    S.slot_state[j] = ServeState::kStaged;
    S.slot_off[j] = off;
    S.add_rows(r, +1);
    S.slot_req[j] = std::move(r);  // its host vectors stay alive until the slot retires
    S.staged_pending = true;
    int64_t f = 0;
    for (const auto& kv : S.free_ranges) f += kv.second;
    S.peak_in_use = std::max(S.peak_in_use, S.pool - f);
    return true;
}

// Start the columns of every staged request: the compute stream waits for the uploads,
// then each column's x = 0, r = b, z = Q^-1 r, p = z and its recurrence scalars.
void admit_staged(globe_ctx* c, ServeState& S) {
    if (!S.staged_pending) return;
    hipStream_t s = c->stream;
    HIP_CHECK(hipEventRecord(S.staged_ev, S.copy));
    HIP_CHECK(hipStreamWaitEvent(s, S.staged_ev, 0));
    for (int j = 0; j < S.slots; ++j) {
        if (S.slot_state[j] != ServeState::kStaged) continue;
        const ServeState::Request& r = S.slot_req[j];
        const int64_t off = S.slot_off[j], nrec = r.nrec;
        launch_set_column(s, j, off, r.ntiles, nrec, r.theta, S.col_off.as<int64_t>(), S.ntiles_col.as<int32_t>(),
                          S.nrec_col.as<int64_t>(), S.theta.as<double>());
        double* b = S.B.as<double>() + off;
        double* x = S.X.as<double>() + off;
        double* rr = S.R.as<double>() + off;
        double* z = S.Z.as<double>() + off;
        double* p = S.P.as<double>() + off;
        launch_fill(s, x, nrec, 0.0);
        launch_copy_cols(s, nrec, b, nrec, rr, nrec, 1);
        // z = Q^-1 r with the partial r.z and r.r; |b| partials; p = z; then the scalars
        const RecMeta m = S.meta(j);
        launch_precond(s, m, c->d_gdiag.as<double>(), S.theta.as<double>() + 5 * j, S.R.as<double>(), 0, 1,
                       S.Z.as<double>(), 0, S.part_c.as<double>(), S.part_b.as<double>(), nullptr, 0);
        launch_dot_parts(s, nrec, b, nrec, b, nrec, 1, S.part_a.as<double>());
        launch_copy_cols(s, nrec, z, nrec, p, nrec, 1);
        launch_pcg_admit(s, S.state.as<PcgState>(), j, S.part_a.as<double>(), parts_for(nrec),
                         S.part_b.as<double>(), S.part_c.as<double>(), precond_parts(m));
        S.slot_state[j] = ServeState::kBusy;
        S.admit_iter[j] = S.global_iter;
        S.t_admit[j] = S.now();
        S.rows_dirty = true;
    }
    S.staged_pending = false;
}

// The individuals the block's genotype passes need (those its requests hold), when they
// are at most 80% of the cohort; otherwise every individual. The listed rows stream at
// about 5.0-5.4 TB/s against 5.9-6.0 for a full pass, so reading fewer pays up to about
// 85%. Called with nothing in flight on the stream that reads rows_dev.
void update_rows(globe_ctx* c, ServeState& S) {
    S.rows_dirty = false;
    S.rows_on = c->gather_on;
    S.rows_n = 0;
    if (!c->gather_on || !S.count_rows || c->gen != 1) return;
    const int32_t n = c->n;
    const int64_t cnt = S.union_cnt;
    const int64_t ng = round_up(cnt, kRowPad);
    if (cnt == 0 || 10 * cnt > 8 * (int64_t)n || ng > n) return;
    S.rows_host.clear();
    S.rows_host.reserve((size_t)ng);
    const int32_t* refs = S.ind_refs.data();
    for (int32_t i = 0; i < n; ++i)
        if (refs[i] > 0) S.rows_host.push_back(i);
    for (int32_t i = 0; i < n && (int64_t)S.rows_host.size() < ng; ++i)
        if (refs[i] == 0) S.rows_host.push_back(i);
    HIP_CHECK(hipMemcpyAsync(S.rows_dev.p, S.rows_host.data(), sizeof(int32_t) * (size_t)ng,
                             hipMemcpyHostToDevice, c->stream));
    S.rows_n = ng;
}

// one iteration of the block PCG over columns [0, k)
void iterate(globe_ctx* c, ServeState& S, int k) {
    hipStream_t s = c->stream;
    const int32_t n = c->n;
    const int64_t cap = S.cap;
    PcgState* st = S.state.as<PcgState>();
    const int* skip = &st->done;
    double* S0 = ws<double>(c, c->w_S0, sizeof(double) * (size_t)n * k);
    double* S1 = ws<double>(c, c->w_S1, sizeof(double) * (size_t)n * k);
    double* GZ = ws<double>(c, c->w_GZ, sizeof(double) * (size_t)n * k);
    const RecMeta m = S.meta(0);
    const double* th = S.theta.as<double>();
    double* part_a = S.part_a.as<double>();
    double* part_b = S.part_b.as<double>();
    double* part_c = S.part_c.as<double>();
    // V p: individuals outside a request's subset keep S0 = S1 = 0
    launch_fill(s, S0, (int64_t)n * k, 0.0);
    launch_fill(s, S1, (int64_t)n * k, 0.0);
    // p = z + beta p of the active slots fused into the sums (an admitted slot starts
    // with p = z and beta = 0), alpha with the x/r update and the preconditioner, and the
    // convergence test with beta in one block (the slots' records are always tiled)
    launch_seg_sums_pcg(s, m, S.P.as<double>(), S.Z.as<double>(), 0, k, st, S0, S1, n);
    {
        // the individual-major pass forms only the requests' individuals' rows
        struct RowsGuard {
            globe_ctx* c;
            ~RowsGuard() {
                c->gather_idx = nullptr;
                c->gather_n = 0;
            }
        } rows_guard{c};
        c->gather_idx = S.rows_dev.as<int32_t>();
        c->gather_n = S.rows_n;
        genetic_product(c, S0, k, GZ, S.o.precision, skip, th);
    }
    launch_record_out(s, m, S.P.as<double>(), 0, k, GZ, S0, S1, n, th, S.VP.as<double>(), 0, part_a, skip);
    launch_pcg_xr_precond(s, m, c->d_gdiag.as<double>(), th, st, part_a, parts_for(cap), S.X.as<double>(), 0,
                          S.R.as<double>(), S.P.as<double>(), S.VP.as<double>(), 0, k, S.Z.as<double>(), part_c,
                          part_b);
    launch_pcg_check_beta(s, st, part_b, precond_parts(m), part_c, precond_parts(m), k);
    ++S.global_iter;
}

}  // namespace

void serve_release(globe_ctx* c) {
    if (!c->serve) return;
    ServeState* S = c->serve;
    if (S->copy) (void)hipStreamSynchronize(S->copy);
    if (S->staged_ev) (void)hipEventDestroy(S->staged_ev);
    if (S->copy) (void)hipStreamDestroy(S->copy);
    if (S->h_state) (void)hipHostFree(S->h_state);
    delete S;
    c->serve = nullptr;
}

}  // namespace globe

using namespace globe;

extern "C" {

int globe_serve_open(globe_ctx* c, const globe_serve_opts* o) {
    API_BEGIN(c)
    require(o != nullptr, GLOBE_E_INVALID, "null options");
    require(c->has_cohort && c->gen != 0, GLOBE_E_STATE, "cohort and genetics required");
    require(o->slots >= 1 && o->slots <= kMaxCols, GLOBE_E_INVALID, "slots must be in [1, 64]");
    require(o->max_records >= 1, GLOBE_E_INVALID, "max_records must be positive");
    require(o->hbm_budget >= 0, GLOBE_E_INVALID, "hbm_budget must be >= 0");
    require(o->tol > 0.0 && o->max_iter >= 1, GLOBE_E_INVALID, "tol and max_iter must be positive");
    require(o->precision >= GLOBE_PREC_FP64 && o->precision <= GLOBE_PREC_FP64_VALU, GLOBE_E_INVALID,
            "unknown precision");
    serve_release(c);
    auto* S = new ServeState();
    c->serve = S;
    S->o = *o;
    S->slots = o->slots;
    S->cap = o->max_records;
    S->pool = o->hbm_budget > 0 ? o->hbm_budget / kPoolElemBytes : request_span(S->cap) * S->slots;
    require(S->pool >= 64, GLOBE_E_INVALID, "hbm_budget holds no request");
    S->free_ranges.emplace(0, S->pool);
    const size_t pool = (size_t)S->pool;
    for (DevBuf* d : {&S->B, &S->X, &S->R, &S->Z, &S->P, &S->VP}) d->ensure(sizeof(double) * pool);
    S->indiv.ensure(sizeof(int32_t) * pool);
    S->times.ensure(sizeof(double) * pool);
    S->rec_blk.ensure(sizeof(int32_t) * pool);
    S->blk_off.ensure(sizeof(int64_t) * pool);
    S->blk_ind.ensure(sizeof(int32_t) * pool);
    S->tile_blk.ensure(sizeof(int32_t) * pool);
    S->col_off.ensure(sizeof(int64_t) * kMaxCols);
    S->ntiles_col.ensure(sizeof(int32_t) * kMaxCols);
    S->nrec_col.ensure(sizeof(int64_t) * kMaxCols);
    S->theta.ensure(sizeof(double) * 5 * kMaxCols);
    S->state.ensure(sizeof(PcgState));
    for (DevBuf* d : {&S->part_a, &S->part_b, &S->part_c}) d->ensure(sizeof(double) * kMaxParts * kMaxCols);
    S->count_rows = c->gen == 1;
    if (S->count_rows) {
        S->ind_refs.assign((size_t)c->n, 0);
        S->rows_dev.ensure(sizeof(int32_t) * ((size_t)c->n + kRowPad));
    }
    hipStream_t s = c->stream;
    HIP_CHECK(hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&S->staged_ev, hipEventDisableTiming));
    // free slots: no records, harmless components
    const double th_free[5] = {1.0, 1.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < kMaxCols; ++j)
        launch_set_column(s, j, 0, 0, 0, th_free, S->col_off.as<int64_t>(), S->ntiles_col.as<int32_t>(),
                          S->nrec_col.as<int64_t>(), S->theta.as<double>());
    HIP_CHECK(hipHostMalloc((void**)&S->h_state, sizeof(PcgState), hipHostMallocDefault));
    std::memset(S->h_state, 0, sizeof(PcgState));
    S->h_state->max_iter = INT_MAX;
    S->h_state->rule = GLOBE_RULE_PER_COLUMN;
    S->h_state->k = S->slots;
    S->h_state->tol = o->tol;
    S->h_state->done = 1;
    HIP_CHECK(hipMemcpyAsync(S->state.p, S->h_state, sizeof(PcgState), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    S->slot_req.resize(S->slots);
    S->slot_state.assign(S->slots, ServeState::kFree);
    S->slot_off.assign(S->slots, 0);
    S->admit_iter.assign(S->slots, 0);
    S->t_admit.assign(S->slots, 0.0);
    ensure_gdiag(c);
    // tune the contraction for every width the block can take before requests arrive
    pretune_geno_product(c, S->slots, o->precision, S->theta.as<double>());
    S->t0 = std::chrono::steady_clock::now();
    API_END(c)
}

int globe_serve_submit(globe_ctx* c, const globe_request* q, int64_t* id) {
    API_BEGIN(c)
    ServeState& S = server(c);
    ServeState::Request r;
    prepare(c, S, q, r);
    require(r.span <= S.pool, GLOBE_E_INVALID, "request larger than the HBM budget");
    std::lock_guard<std::mutex> lk(S.qmu);
    r.id = S.next_id++;
    r.t_submit = S.now();
    if (id) *id = r.id;
    S.queue.push_back(std::move(r));
    API_END(c)
}

int globe_serve_step(globe_ctx* c, int32_t* finished, int32_t* active, int32_t* queued) {
    API_BEGIN(c)
    ServeState& S = server(c);
    hipStream_t s = c->stream;
    int fin = 0;
    // The iteration the previous step enqueued: wait for it, then retire the columns its
    // state shows finished (converged, or at the iteration cap). The step returns without
    // waiting for its own iteration, so the caller's work between steps (taking results,
    // its intake) overlaps the GPU. A retired column's records leave the block before the
    // next iteration; its solution no longer changes, so it is copied out on the copy
    // stream while that iteration runs, and only then does its slot and range free up.
    std::vector<ServeState::Retire> retiring;
    if (S.inflight) {
        S.inflight = false;
        HIP_CHECK(hipStreamSynchronize(s));
        const double t = S.now();
        const double th_free[5] = {1.0, 1.0, 0.0, 0.0, 0.0};
        for (int j = 0; j < S.slots; ++j) {
            if (S.slot_state[j] != ServeState::kBusy) continue;
            const int its = S.global_iter - S.admit_iter[j];
            const bool conv = S.h_state->active[j] == 0;
            if (!conv && its < S.o.max_iter) continue;
            const ServeState::Request& r = S.slot_req[j];
            ServeState::Retire q;
            q.slot = j;
            q.res.id = r.id;
            q.res.iters = its;
            q.res.converged = conv ? 1 : 0;
            q.res.ratio = S.h_state->ratio[j];
            q.res.t_submit = r.t_submit;
            q.res.t_admit = S.t_admit[j];
            q.res.t_done = t;
            // the column holds no records and does no further work
            launch_set_column(s, j, 0, 0, 0, th_free, S.col_off.as<int64_t>(), S.ntiles_col.as<int32_t>(),
                              S.nrec_col.as<int64_t>(), S.theta.as<double>());
            if (!conv) launch_pcg_deactivate(s, S.state.as<PcgState>(), j);
            S.slot_state[j] = ServeState::kRetiring;
            retiring.push_back(q);
        }
    }
    auto busy_cols = [&]() {
        int k = 0;
        for (int j = 0; j < S.slots; ++j)
            if (S.slot_state[j] == ServeState::kBusy) k = j + 1;
        return k;
    };
    // an idle block takes the queue's head at once; a running one joins the requests
    // staged during the previous iteration
    if (busy_cols() == 0 && !S.staged_pending)
        while (stage_next(S)) {
        }
    admit_staged(c, S);
    if (S.rows_dirty || S.rows_on != c->gather_on) update_rows(c, S);
    const int k = busy_cols();
    if (k > 0) {
        // mixed precision: each slot's digits from its own residual ratio (4 in its first
        // iteration), none for a slot that holds no request (slot k - 1 is busy, so the
        // contraction keeps at least one digit column)
        struct LimbsGuard {
            globe_ctx* c;
            ~LimbsGuard() {
                c->limbs_col.clear();
                c->limbs_mixed = c->limbs_default;
            }
        } guard{c};
        if (S.o.precision == GLOBE_PREC_MIXED) {
            c->limbs_col.assign((size_t)k, (int8_t)0);
            int lmax = 2;
            for (int j = 0; j < k; ++j) {
                if (S.slot_state[j] != ServeState::kBusy) continue;
                const int L = S.global_iter == S.admit_iter[j] ? kLimbsMixed
                                                               : relaxed_limbs(S.h_state->ratio[j], S.o.tol);
                c->limbs_col[j] = (int8_t)L;
                lmax = std::max(lmax, L);
            }
            c->limbs_mixed = lmax;
        }
        iterate(c, S, k);
        HIP_CHECK(hipMemcpyAsync(S.h_state, S.state.p, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        S.inflight = true;
    }
    // the retired columns' solutions (the iteration just enqueued writes none of their
    // elements); their ranges return to the pool once the copies are done, ahead of any
    // upload into them on the same stream
    if (!retiring.empty()) {
        for (const ServeState::Retire& q : retiring) {
            const ServeState::Request& r = S.slot_req[q.slot];
            ServeState::Done d;
            d.x.resize(r.nrec);
            HIP_CHECK(hipMemcpyAsync(d.x.data(), S.X.as<double>() + S.slot_off[q.slot], sizeof(double) * r.nrec,
                                     hipMemcpyDeviceToHost, S.copy));
            d.res = q.res;
            S.done.emplace(q.res.id, std::move(d));
        }
        HIP_CHECK(hipStreamSynchronize(S.copy));
        for (const ServeState::Retire& q : retiring) {
            const int j = q.slot;
            S.release(S.slot_off[j], S.slot_req[j].span);
            S.add_rows(S.slot_req[j], -1);
            S.slot_state[j] = ServeState::kFree;
            S.slot_req[j] = ServeState::Request();
            ++fin;
        }
    }
    // the next requests upload while the iteration runs
    while (stage_next(S)) {
    }
    if (finished) *finished = fin;
    if (active) {
        int a = 0;
        for (int j = 0; j < S.slots; ++j) a += S.slot_state[j] == ServeState::kBusy ? 1 : 0;
        *active = a;
    }
    if (queued) {  // waiting: in the queue, or staged to join at the next step
        int q = 0;
        for (int j = 0; j < S.slots; ++j) q += S.slot_state[j] == ServeState::kStaged ? 1 : 0;
        std::lock_guard<std::mutex> lk(S.qmu);
        *queued = (int32_t)S.queue.size() + q;
    }
    API_END(c)
}

int globe_serve_finished(globe_ctx* c, int64_t* ids, int32_t max, int32_t* count) {
    API_BEGIN(c)
    ServeState& S = server(c);
    int32_t n = 0;
    for (const auto& kv : S.done) {
        if (n >= max) break;
        if (ids) ids[n] = kv.first;
        ++n;
    }
    if (count) *count = n;
    API_END(c)
}

int globe_serve_take(globe_ctx* c, int64_t id, double* x, globe_serve_result* res) {
    API_BEGIN(c)
    ServeState& S = server(c);
    auto it = S.done.find(id);
    require(it != S.done.end(), GLOBE_E_STATE, "request " + std::to_string(id) + " has not finished");
    if (x) std::memcpy(x, it->second.x.data(), sizeof(double) * it->second.x.size());
    if (res) *res = it->second.res;
    S.done.erase(it);
    API_END(c)
}

int globe_serve_usage(globe_ctx* c, int64_t out[4]) {
    API_BEGIN(c)
    ServeState& S = server(c);
    require(out != nullptr, GLOBE_E_INVALID, "null output");
    int64_t f = 0;
    for (const auto& kv : S.free_ranges) f += kv.second;
    out[0] = (S.pool - f) * kPoolElemBytes;
    out[1] = S.peak_in_use * kPoolElemBytes;
    out[2] = S.pool * kPoolElemBytes;
    out[3] = S.waits_for_room;
    API_END(c)
}

int globe_serve_close(globe_ctx* c) {
    API_BEGIN(c)
    HIP_CHECK(hipStreamSynchronize(c->stream));
    serve_release(c);
    API_END(c)
}

}  // extern "C"
