// Serving loop: one resident cohort, solve requests admitted into the columns of one
// shared block PCG as slots and HBM free up (continuous batching).
//
// Slot j owns column j of every record-space matrix ([slots][cap] doubles: b, x, r,
// z, p, V p) and its own record structure at the same column stride (individual and
// time per record, individual blocks, LDS tiles) -- the kernels read it through a
// RecMeta with per-column strides and counts. Records past a request's own count are
// zero in b, x, r, z, p and V p, so the column-wise PCG kernels run over the full
// capacity unchanged. The genotype contraction takes all occupied columns at once;
// its fixed-point scale, the PCG scalars and the record structure are per column, so
// a request's arithmetic does not depend on which other requests share the block.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <deque>
#include <map>

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
        int64_t bytes = 0;
    };
    struct Done {
        std::vector<double> x;
        globe_serve_result res{};
    };
    globe_serve_opts o{};
    int slots = 0;
    int64_t cap = 0;
    globe::DevBuf B, X, R, Z, P, VP;                           // [slots][cap] doubles
    // per-slot structure: per record [slots][cap], per block and per tile [slots][cap + 1]
    // (blk_off, blk_ind and tile_blk share the block stride blk_ld = tile_ld = cap + 1)
    globe::DevBuf indiv, times, rec_blk, blk_off, blk_ind, tile_blk;
    globe::DevBuf ntiles_col, nrec_col, theta, state;
    globe::DevBuf part_a, part_b, part_c;
    std::deque<Request> queue;
    std::vector<Request> slot_req;
    std::vector<char> busy;
    std::vector<int> admit_iter;
    std::vector<double> t_admit;
    std::map<int64_t, Done> done;
    int64_t next_id = 0, bytes_in_use = 0, budget = 0;
    int global_iter = 0;
    std::chrono::steady_clock::time_point t0;
    globe::PcgState* h_state = nullptr;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    // record structure of slots [j0, j0 + k)
    globe::RecMeta meta(int j0) const {
        globe::RecMeta m;
        m.indiv = indiv.as<int32_t>() + j0 * cap;
        m.times = times.as<double>() + j0 * cap;
        m.rec_blk = rec_blk.as<int32_t>() + j0 * cap;
        m.rec_ld = cap;
        m.blk_off = blk_off.as<int64_t>() + j0 * (cap + 1);
        m.blk_ind = blk_ind.as<int32_t>() + j0 * (cap + 1);
        m.blk_ld = cap + 1;
        m.tile_blk = tile_blk.as<int32_t>() + j0 * (cap + 1);
        m.tile_ld = cap + 1;
        m.ntiles_col = ntiles_col.as<int32_t>() + j0;
        // two consecutive tiles hold more than kTileRecords records, so a column has
        // at most 2 cap / kTileRecords + 1 tiles (the launch grid covers that many)
        m.ntiles = (int32_t)(2 * cap / globe::kTileRecords + 1);
        m.nrec_col = nrec_col.as<int64_t>() + j0;
        m.nrec = cap;
        m.cap = globe::kTileRecords;
        return m;
    }
};

namespace globe {

namespace {

ServeState& server(globe_ctx* c) {
    require(c->serve != nullptr, GLOBE_E_STATE, "serving loop not open (globe_serve_open)");
    return *c->serve;
}

// bytes of device memory a request occupies while admitted: six record vectors and
// its record structure
int64_t request_bytes(int64_t nrec) { return nrec * (6 * 8 + 4 + 8 + 4 + 8 + 4 + 4); }

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
        if (in_tile > 0 && in_tile + len > kTileRecords) {
            r.tile_blk.push_back(q2);
            in_tile = 0;
        }
        in_tile += len;
    }
    r.tile_blk.push_back(r.nblk);
    r.ntiles = (int32_t)r.tile_blk.size() - 1;
    r.bytes = request_bytes(nrec);
}

// place request r in free slot j and start its column
void admit(globe_ctx* c, ServeState& S, int j, ServeState::Request&& r) {
    hipStream_t s = c->stream;
    const int64_t cap = S.cap, nrec = r.nrec;
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    };
    h2d(S.indiv.as<int32_t>() + j * cap, r.indiv.data(), sizeof(int32_t) * nrec);
    h2d(S.times.as<double>() + j * cap, r.times.data(), sizeof(double) * nrec);
    h2d(S.rec_blk.as<int32_t>() + j * cap, r.rec_blk.data(), sizeof(int32_t) * nrec);
    h2d(S.blk_off.as<int64_t>() + j * (cap + 1), r.blk_off.data(), sizeof(int64_t) * r.blk_off.size());
    h2d(S.blk_ind.as<int32_t>() + j * (cap + 1), r.blk_ind.data(), sizeof(int32_t) * r.blk_ind.size());
    h2d(S.tile_blk.as<int32_t>() + j * (cap + 1), r.tile_blk.data(), sizeof(int32_t) * r.tile_blk.size());
    h2d(S.ntiles_col.as<int32_t>() + j, &r.ntiles, sizeof(int32_t));
    h2d(S.nrec_col.as<int64_t>() + j, &r.nrec, sizeof(int64_t));
    h2d(S.theta.as<double>() + 5 * j, r.theta, sizeof(double) * 5);
    double* b = S.B.as<double>() + j * cap;
    double* x = S.X.as<double>() + j * cap;
    double* rr = S.R.as<double>() + j * cap;
    double* z = S.Z.as<double>() + j * cap;
    double* p = S.P.as<double>() + j * cap;
    h2d(b, r.y.data(), sizeof(double) * nrec);
    if (nrec < cap) {
        launch_fill(s, b + nrec, cap - nrec, 0.0);
        launch_fill(s, z + nrec, cap - nrec, 0.0);
    }
    launch_fill(s, x, cap, 0.0);
    launch_copy_cols(s, cap, b, cap, rr, cap, 1);
    // z = Q^-1 r, partial r.z and r.r; |b| partials; p = z; then the column's scalars
    const RecMeta m = S.meta(j);
    launch_precond(s, m, c->d_gdiag.as<double>(), S.theta.as<double>() + 5 * j, rr, cap, 1, z, cap,
                   S.part_c.as<double>(), S.part_b.as<double>(), nullptr, 0);
    launch_dot_parts(s, cap, b, cap, b, cap, 1, S.part_a.as<double>());
    launch_copy_cols(s, cap, z, cap, p, cap, 1);
    launch_pcg_admit(s, S.state.as<PcgState>(), j, S.part_a.as<double>(), parts_for(cap),
                     S.part_b.as<double>(), S.part_c.as<double>(), precond_parts(m));
    // the copies read the request's host vectors: finish them before r is moved
    HIP_CHECK(hipStreamSynchronize(s));
    S.bytes_in_use += r.bytes;
    S.busy[j] = 1;
    S.admit_iter[j] = S.global_iter;
    S.t_admit[j] = S.now();
    S.slot_req[j] = std::move(r);
}

// one iteration of the block PCG over slots [0, k)
void iterate(globe_ctx* c, ServeState& S, int k) {
    hipStream_t s = c->stream;
    const int32_t n = c->n;
    const int64_t cap = S.cap;
    PcgState* st = S.state.as<PcgState>();
    const int* skip = &st->done;
    c->w_S0.ensure(sizeof(double) * (size_t)n * k);
    c->w_S1.ensure(sizeof(double) * (size_t)n * k);
    c->w_GZ.ensure(sizeof(double) * (size_t)n * k);
    double* S0 = c->w_S0.as<double>();
    double* S1 = c->w_S1.as<double>();
    double* GZ = c->w_GZ.as<double>();
    const RecMeta m = S.meta(0);
    const double* th = S.theta.as<double>();
    double* part_a = S.part_a.as<double>();
    double* part_b = S.part_b.as<double>();
    double* part_c = S.part_c.as<double>();
    // V p: individuals outside a request's subset keep S0 = S1 = 0
    launch_fill(s, S0, (int64_t)n * k, 0.0);
    launch_fill(s, S1, (int64_t)n * k, 0.0);
    launch_seg_sums(s, m, S.P.as<double>(), cap, k, S0, S1, n, skip);
    genetic_product(c, S0, k, GZ, S.o.precision, skip, th);
    launch_record_out(s, m, S.P.as<double>(), cap, k, GZ, S0, S1, n, th, S.VP.as<double>(), cap, part_a,
                      skip, st);
    launch_pcg_xr(s, st, cap, S.X.as<double>(), cap, S.R.as<double>(), S.P.as<double>(),
                  S.VP.as<double>(), cap, k, part_b);
    launch_precond(s, m, c->d_gdiag.as<double>(), th, S.R.as<double>(), cap, k, S.Z.as<double>(), cap,
                   part_c, nullptr, st, 0, true);
    launch_pcg_p(s, st, cap, S.P.as<double>(), S.Z.as<double>(), cap, k);
    ++S.global_iter;
}

}  // namespace

void serve_release(globe_ctx* c) {
    if (!c->serve) return;
    if (c->serve->h_state) (void)hipHostFree(c->serve->h_state);
    delete c->serve;
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
    require(o->tol > 0.0 && o->max_iter >= 1, GLOBE_E_INVALID, "tol and max_iter must be positive");
    require(o->precision >= GLOBE_PREC_FP64 && o->precision <= GLOBE_PREC_FP64_VALU, GLOBE_E_INVALID,
            "unknown precision");
    serve_release(c);
    auto* S = new ServeState();
    c->serve = S;
    S->o = *o;
    S->slots = o->slots;
    S->cap = round_up(o->max_records, 64);
    S->budget = o->hbm_budget > 0 ? o->hbm_budget : request_bytes(S->cap) * S->slots;
    const size_t vec = sizeof(double) * (size_t)S->cap * S->slots;
    for (DevBuf* d : {&S->B, &S->X, &S->R, &S->Z, &S->P, &S->VP}) d->ensure(vec);
    S->indiv.ensure(sizeof(int32_t) * (size_t)S->cap * S->slots);
    S->times.ensure(sizeof(double) * (size_t)S->cap * S->slots);
    S->rec_blk.ensure(sizeof(int32_t) * (size_t)S->cap * S->slots);
    S->blk_off.ensure(sizeof(int64_t) * (size_t)(S->cap + 1) * S->slots);
    S->blk_ind.ensure(sizeof(int32_t) * (size_t)(S->cap + 1) * S->slots);
    S->tile_blk.ensure(sizeof(int32_t) * (size_t)(S->cap + 1) * S->slots);
    S->ntiles_col.ensure(sizeof(int32_t) * kMaxCols);
    S->nrec_col.ensure(sizeof(int64_t) * kMaxCols);
    S->theta.ensure(sizeof(double) * 5 * kMaxCols);
    S->state.ensure(sizeof(PcgState));
    for (DevBuf* d : {&S->part_a, &S->part_b, &S->part_c}) d->ensure(sizeof(double) * kMaxParts * kMaxCols);
    hipStream_t s = c->stream;
    for (DevBuf* d : {&S->B, &S->X, &S->R, &S->Z, &S->P, &S->VP})
        launch_fill(s, d->as<double>(), (int64_t)S->cap * S->slots, 0.0);
    HIP_CHECK(hipMemsetAsync(S->ntiles_col.p, 0, sizeof(int32_t) * kMaxCols, s));
    HIP_CHECK(hipMemsetAsync(S->nrec_col.p, 0, sizeof(int64_t) * kMaxCols, s));
    // free slots carry harmless components until a request takes them
    std::vector<double> th((size_t)5 * kMaxCols);
    for (int j = 0; j < kMaxCols; ++j) {
        th[5 * j + 0] = 1.0;
        th[5 * j + 1] = 1.0;
    }
    HIP_CHECK(hipMemcpyAsync(S->theta.p, th.data(), th.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipHostMalloc((void**)&S->h_state, sizeof(PcgState), hipHostMallocDefault));
    std::memset(S->h_state, 0, sizeof(PcgState));
    S->h_state->max_iter = INT_MAX;
    S->h_state->rule = GLOBE_RULE_PER_COLUMN;
    S->h_state->k = S->slots;
    S->h_state->tol = o->tol;
    S->h_state->done = 1;
    HIP_CHECK(hipMemcpyAsync(S->state.p, S->h_state, sizeof(PcgState), hipMemcpyHostToDevice, s));
    S->slot_req.resize(S->slots);
    S->busy.assign(S->slots, 0);
    S->admit_iter.assign(S->slots, 0);
    S->t_admit.assign(S->slots, 0.0);
    ensure_gdiag(c);
    // tune the contraction for every width the block can take before requests arrive
    c->w_S0.ensure(sizeof(double) * (size_t)c->n * S->slots);
    c->w_GZ.ensure(sizeof(double) * (size_t)c->n * S->slots);
    launch_fill(s, c->w_S0.as<double>(), (int64_t)c->n * S->slots, 0.0);
    for (int k : {8, 16, 32, 64}) {
        const int kk = std::min(k, S->slots);
        genetic_product(c, c->w_S0.as<double>(), kk, c->w_GZ.as<double>(), o->precision, nullptr,
                        S->theta.as<double>());
        if (kk == S->slots) break;
    }
    HIP_CHECK(hipStreamSynchronize(s));
    S->t0 = std::chrono::steady_clock::now();
    API_END(c)
}

int globe_serve_submit(globe_ctx* c, const globe_request* q, int64_t* id) {
    API_BEGIN(c)
    ServeState& S = server(c);
    ServeState::Request r;
    prepare(c, S, q, r);
    require(r.bytes <= S.budget, GLOBE_E_INVALID, "request larger than the HBM budget");
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
    // admit FIFO while a slot is free and the budget holds
    while (!S.queue.empty()) {
        int j = 0;
        while (j < S.slots && S.busy[j]) ++j;
        if (j == S.slots || S.bytes_in_use + S.queue.front().bytes > S.budget) break;
        ServeState::Request r = std::move(S.queue.front());
        S.queue.pop_front();
        admit(c, S, j, std::move(r));
    }
    int k = 0;
    for (int j = 0; j < S.slots; ++j)
        if (S.busy[j]) k = j + 1;
    int fin = 0;
    if (k > 0) {
        // mixed precision: each slot's digits from its own residual ratio (4 in its first
        // iteration), one digit column of zeros for an empty slot
        struct LimbsGuard {
            globe_ctx* c;
            ~LimbsGuard() {
                c->limbs_col.clear();
                c->limbs_mixed = c->limbs_default;
            }
        } guard{c};
        if (S.o.precision == GLOBE_PREC_MIXED) {
            c->limbs_col.assign((size_t)k, (int8_t)1);
            int lmax = 2;
            for (int j = 0; j < k; ++j) {
                if (!S.busy[j]) continue;
                const int L = S.global_iter == S.admit_iter[j] ? kLimbsMixed
                                                               : relaxed_limbs(S.h_state->ratio[j], S.o.tol);
                c->limbs_col[j] = (int8_t)L;
                lmax = std::max(lmax, L);
            }
            c->limbs_mixed = lmax;
        }
        iterate(c, S, k);
        HIP_CHECK(hipMemcpyAsync(S.h_state, S.state.p, sizeof(PcgState), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        const double t = S.now();
        for (int j = 0; j < k; ++j) {
            if (!S.busy[j]) continue;
            const int its = S.global_iter - S.admit_iter[j];
            const bool conv = S.h_state->active[j] == 0;
            if (!conv && its < S.o.max_iter) continue;
            ServeState::Request& r = S.slot_req[j];
            ServeState::Done d;
            d.x.resize(r.nrec);
            HIP_CHECK(hipMemcpyAsync(d.x.data(), S.X.as<double>() + j * S.cap, sizeof(double) * r.nrec,
                                     hipMemcpyDeviceToHost, s));
            d.res.id = r.id;
            d.res.iters = its;
            d.res.converged = conv ? 1 : 0;
            d.res.ratio = S.h_state->ratio[j];
            d.res.t_submit = r.t_submit;
            d.res.t_admit = S.t_admit[j];
            d.res.t_done = t;
            // retire the column: it stops contributing and takes no further work
            int32_t zero32 = 0;
            int64_t zero64 = 0;
            HIP_CHECK(hipMemcpyAsync(S.ntiles_col.as<int32_t>() + j, &zero32, sizeof(int32_t),
                                     hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(S.nrec_col.as<int64_t>() + j, &zero64, sizeof(int64_t),
                                     hipMemcpyHostToDevice, s));
            launch_fill(s, S.P.as<double>() + j * S.cap, S.cap, 0.0);
            if (!conv) {
                int32_t off = 0;
                HIP_CHECK(hipMemcpyAsync(&S.state.as<PcgState>()->active[j], &off, sizeof(int32_t),
                                         hipMemcpyHostToDevice, s));
            }
            HIP_CHECK(hipStreamSynchronize(s));
            S.bytes_in_use -= r.bytes;
            S.busy[j] = 0;
            S.done.emplace(r.id, std::move(d));
            r = ServeState::Request();
            ++fin;
        }
    }
    if (finished) *finished = fin;
    if (active) {
        int a = 0;
        for (int j = 0; j < S.slots; ++j) a += S.busy[j] ? 1 : 0;
        *active = a;
    }
    if (queued) *queued = (int32_t)S.queue.size();
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

int globe_serve_close(globe_ctx* c) {
    API_BEGIN(c)
    HIP_CHECK(hipStreamSynchronize(c->stream));
    serve_release(c);
    API_END(c)
}

}  // extern "C"
