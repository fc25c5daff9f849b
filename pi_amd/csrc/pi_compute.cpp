// pi to D decimal digits by the Chudnovsky series, binary splitting with the
// tree's large merges on the GPU.
//
// Restates pi-core's Chudnovsky (crates/pi-core/src/algorithm/chudnovsky.rs):
//   n_terms = ceil(D / 14.181647462725478) + 8                 (:50, :91)
//   leaf k: p = -(6k-5)(2k-1)(6k-1), q = k^3 C^3/24, t = p (A + B k)   (:379-393)
//   merge:  t = t_l q_r + p_l t_r, p = p_l p_r, q = q_l q_r   (:411-413)
//   root:   p skipped                                          (:339-369)
//   pi = 426880 sqrt(10005) Q / (A Q + T)                      (:146-211)
// at prec = ceil(D log2 10) + 256 bits (precision.rs:19, :64), truncated to D
// digits (util.rs:25-83) and written as "3." + (D-1) digits + "\n"
// (output.rs:33-61).
//
// The split tree is the reference's (m = (a+b)/2).  Nodes whose smaller
// operand is under a limb threshold run whole on host threads (GMP, as the
// reference runs its small products on the CPU); every node above is merged
// on the GPU, a tree level at a time, through the serving loop: all products
// of a level are submitted together (continuous batching groups them into
// shared launches), each t = x +- y is submitted as soon as its two products
// finish, and the level's P, Q, T stay in HBM as the next level's operands.
//
// Sign: T(a,b) < 0 for every range (an alternating series whose terms shrink,
// led by the negative k = a term), so the device holds |T| and
//   |T| = |T_l| Q_r + (-1)^(m-a) |P_l| |T_r|,
// a sum when the left half has an even number of terms and otherwise a
// difference that is positive by the same argument.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gnt.h"
#include "gmp_abi.hpp"
#include "par_mpz.hpp"

namespace {

constexpr double DIGITS_PER_TERM = 14.181647462725478;  // chudnovsky.rs:50
constexpr uint64_t SAFETY_TERMS = 8;                    // chudnovsky.rs:56
constexpr uint64_t SAFETY_BITS = 256;                   // precision.rs:19
constexpr unsigned long CH_A = 13591409, CH_B = 545140134;
constexpr unsigned long C3_OVER_24 = 10939058860032000ul;

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

[[noreturn]] void die(const char *what, const char *detail = "") {
    std::fprintf(stderr, "pi_compute: %s %s\n", what, detail);
    std::exit(1);
}

void check(int rc, const char *what) {
    if (rc != GNT_OK) die(what, gnt_last_error());
}

// Serial binary splitting of [a, b) on the host (chudnovsky.rs:377-417).
void bs_host(uint64_t a, uint64_t b, gmp_mpz *P, gmp_mpz *Q, gmp_mpz *T, bool need_p) {
    if (b - a == 1) {
        __gmpz_set_ui(P, 6 * a - 5);
        __gmpz_mul_ui(P, P, 2 * a - 1);
        __gmpz_mul_ui(P, P, 6 * a - 1);
        __gmpz_neg(P, P);
        __gmpz_set_ui(Q, a);
        __gmpz_mul_ui(Q, Q, a);
        __gmpz_mul_ui(Q, Q, a);
        __gmpz_mul_ui(Q, Q, C3_OVER_24);
        __gmpz_mul_ui(T, P, CH_A + CH_B * a);
        return;
    }
    const uint64_t m = (a + b) / 2;
    bs_host(a, m, P, Q, T, true);
    Mpz P2, Q2, T2;
    bs_host(m, b, &P2, &Q2, &T2, need_p);
    __gmpz_mul(T, T, &Q2);
    __gmpz_mul(&T2, P, &T2);
    __gmpz_add(T, T, &T2);
    if (need_p) __gmpz_mul(P, P, &P2);
    __gmpz_mul(Q, Q, &Q2);
}

// log2 |P(a,b)| estimate: sum of log2(72 k^3) over the range.
double p_bits(uint64_t a, uint64_t b) {
    return (double)(b - a) * std::log2(72.0) + 3.0 * (std::lgamma((double)b) - std::lgamma((double)a)) / std::log(2.0);
}

struct Node {
    uint64_t a = 0, b = 0;
    int depth = 0;
    int left = -1, right = -1;  // children of a merged node
    int parent = -1;
    bool host = true;           // computed whole on the host
    std::unique_ptr<Mpz> P, Q, T;
    gnt_buf *bp = nullptr, *bq = nullptr, *bt = nullptr;
};

struct LevelStat {
    int depth;     // a merge wave: the smallest depth in it
    size_t nodes, products;
    double limbs;  // operand limbs (na + nb) summed over the level's products
    double secs;
    int depth_max = -1;  // a merge wave: the largest depth in it (-1: one tree level)
    int wave = -1;
};

struct Options {
    uint64_t digits = 0;
    std::string out;
    double gpu_min_limbs = 1024;  // a node merges on the GPU when both children's P reach this
    int threads = 0;
    bool host_merge = false;      // merge the upper tree on the host too (comparison run)
    bool fa_gmp = false;          // final assembly and radix conversion by GMP (comparison run)
    double fa_min_limbs = 2048;   // final-assembly / conversion products on the GPU from this size
    double dec_min_limbs = 512;   // decimal conversion's batched products on the GPU from this size
                                  // (batches cross as packed transfers: they pay from smaller sizes)
    int device = 0;
    std::string plans;            // plan cache from the autotuner
    std::string algorithm = "chudnovsky";
    double par_min_limbs = 1 << 17;  // host bignum steps multi-threaded from this many limbs
    bool merge_levels = false;    // merge a whole tree level at a time after all host nodes
                                  // (comparison run); default: merge nodes as their children finish
    bool shutdown = false;        // free the context's device memory before exiting
};

struct Driver {
    Options o;
    std::vector<Node> nodes;
    std::vector<int> host_nodes;
    int max_depth = 0;
    gnt_ctx *ctx = nullptr;
    std::vector<LevelStat> levels;

    bool merges_on_device() const { return !o.host_merge; }

    int build(uint64_t a, uint64_t b, int depth) {
        const int id = (int)nodes.size();
        nodes.emplace_back();
        nodes[id].a = a;
        nodes[id].b = b;
        nodes[id].depth = depth;
        max_depth = std::max(max_depth, depth);
        const uint64_t m = (a + b) / 2;
        bool split = b - a >= 2 && std::min(p_bits(a, m), p_bits(m, b)) / 64.0 >= o.gpu_min_limbs;
        if (split) {
            nodes[id].host = false;
            int l = build(a, m, depth + 1);
            int r = build(m, b, depth + 1);
            nodes[id].left = l;
            nodes[id].right = r;
            nodes[l].parent = id;
            nodes[r].parent = id;
        } else {
            host_nodes.push_back(id);
        }
        return id;
    }

    // Host nodes on `threads` threads; with upload set, each node's P, |T|, Q
    // go to the device as soon as it is done (overlapping the other nodes'
    // compute) and the host copies are dropped.
    void run_host_nodes(int threads, bool upload_done = false, bool left_to_right = false) {
        std::vector<int> order = host_nodes;
        if (left_to_right)  // siblings finish together, so their parent can merge early
            std::sort(order.begin(), order.end(), [&](int x, int y) { return nodes[x].a < nodes[y].a; });
        else
            std::sort(order.begin(), order.end(), [&](int x, int y) { return nodes[x].b - nodes[x].a > nodes[y].b - nodes[y].a; });
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= order.size()) return;
                Node &n = nodes[order[i]];
                n.P.reset(new Mpz);
                n.Q.reset(new Mpz);
                n.T.reset(new Mpz);
                bs_host(n.a, n.b, &*n.P, &*n.Q, &*n.T, order[i] != 0);
                if (upload_done && order[i] != 0) {
                    if (n.T->sign() >= 0) die("T(a,b) expected negative", "");
                    n.bp = upload(*n.P);
                    n.bq = upload(*n.Q);
                    n.bt = upload(*n.T);
                    n.P.reset();
                    n.Q.reset();
                    n.T.reset();
                    mark_available(order[i]);
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++) pool.emplace_back(work);
        for (auto &t : pool) t.join();
    }

    gnt_buf *upload(const Mpz &z) {
        gnt_buf *b = nullptr;
        check(gnt_buf_upload(ctx, z.data(), z.limbs(), &b), "upload");
        return b;
    }

    static size_t buf_len(gnt_ctx *c, gnt_buf *b) {
        size_t n = 0;
        check(gnt_buf_len(c, b, &n), "length");
        return n;
    }

    uint64_t submit(int op, gnt_buf *x, gnt_buf *y, double *limbs) {
        gnt_req r{};
        r.op = op;
        r.a_buf = x;
        r.b_buf = y;
        uint64_t t = 0;
        check(gnt_submit(ctx, &r, &t), "submit");
        *limbs += (double)(buf_len(ctx, x) + buf_len(ctx, y));
        return t;
    }

    gnt_buf *wait(uint64_t t) {
        gnt_buf *b = nullptr;
        check(gnt_wait(ctx, t, -1, &b, nullptr), "wait");
        return b;
    }

    // Merge every device node at `depth` (children finished).
    void merge_level_device(int depth) {
        std::vector<int> ids;
        for (int i = 0; i < (int)nodes.size(); i++)
            if (!nodes[i].host && nodes[i].depth == depth) ids.push_back(i);
        if (ids.empty()) return;
        auto t0 = Clock::now();
        LevelStat st{depth, ids.size(), 0, 0.0, 0.0};
        merge_nodes_device(ids, st);
        st.secs = since(t0);
        levels.push_back(st);
    }

    // Merge the given device nodes, whose children are all on the device: all
    // their products submitted together (continuous batching groups them),
    // each t = x +- y as soon as its two products finish.
    void merge_nodes_device(const std::vector<int> &ids, LevelStat &st) {
        struct Tk { uint64_t x, y, q, p, t; };
        std::vector<Tk> tk(ids.size());
        for (size_t i = 0; i < ids.size(); i++) {
            Node &n = nodes[ids[i]];
            Node &L = nodes[n.left], &R = nodes[n.right];
            tk[i].x = submit(GNT_OP_MUL, L.bt, R.bq, &st.limbs);
            tk[i].y = submit(GNT_OP_MUL, L.bp, R.bt, &st.limbs);
            tk[i].q = submit(GNT_OP_MUL, L.bq, R.bq, &st.limbs);
            tk[i].p = ids[i] != 0 ? submit(GNT_OP_MUL, L.bp, R.bp, &st.limbs) : 0;
            st.products += ids[i] != 0 ? 4 : 3;
        }
        std::vector<gnt_buf *> xs(ids.size()), ys(ids.size());
        for (size_t i = 0; i < ids.size(); i++) {
            Node &n = nodes[ids[i]];
            xs[i] = wait(tk[i].x);
            ys[i] = wait(tk[i].y);
            const uint64_t left_terms = nodes[n.left].b - nodes[n.left].a;
            tk[i].t = submit(left_terms % 2 == 0 ? GNT_OP_ADD : GNT_OP_SUB, xs[i], ys[i], &st.limbs);
        }
        for (size_t i = 0; i < ids.size(); i++) {
            Node &n = nodes[ids[i]];
            n.bq = wait(tk[i].q);
            if (ids[i] != 0) n.bp = wait(tk[i].p);
            n.bt = wait(tk[i].t);
            gnt_buf_release(ctx, xs[i]);
            gnt_buf_release(ctx, ys[i]);
            for (int c : {n.left, n.right}) {
                Node &ch = nodes[c];
                gnt_buf_release(ctx, ch.bp);
                gnt_buf_release(ctx, ch.bq);
                gnt_buf_release(ctx, ch.bt);
                ch.bp = ch.bq = ch.bt = nullptr;
            }
        }
    }

    // ---- merges as children finish (overlapping the host nodes) ----
    std::mutex avail_mu;
    std::condition_variable avail_cv;
    std::vector<int> newly_available;  // nodes whose P, Q, |T| just reached the device

    void mark_available(int id) {
        {
            std::lock_guard<std::mutex> g(avail_mu);
            newly_available.push_back(id);
        }
        avail_cv.notify_one();
    }

    // Host nodes on `threads` threads, left to right, while this thread merges
    // every device node whose two children are on the device, in waves of
    // whatever is ready; returns when the root is merged.  *t_host receives
    // the time to the last host node, *t_tail the merging left after it.
    void run_dataflow(int threads, double *t_host, double *t_tail) {
        auto t0 = Clock::now();
        std::atomic<bool> host_done{false};
        Clock::time_point t_host_end;
        std::thread hosts([&] {
            run_host_nodes(threads, true, true);
            t_host_end = Clock::now();
            host_done = true;
        });
        std::vector<char> avail(nodes.size(), 0), queued(nodes.size(), 0);
        size_t left = 0;
        for (const Node &n : nodes) left += n.host ? 0 : 1;
        int wave = 0;
        while (left > 0) {
            std::vector<int> ready;
            {
                std::unique_lock<std::mutex> g(avail_mu);
                avail_cv.wait(g, [&] { return !newly_available.empty(); });
                for (int id : newly_available) {
                    avail[id] = 1;
                    const int p = nodes[id].parent;
                    if (p >= 0 && !queued[p] && avail[nodes[p].left] && avail[nodes[p].right]) {
                        queued[p] = 1;
                        ready.push_back(p);
                    }
                }
                newly_available.clear();
            }
            if (ready.empty()) continue;
            auto tw = Clock::now();
            int dmin = 1 << 30, dmax = -1;
            for (int id : ready) {
                dmin = std::min(dmin, nodes[id].depth);
                dmax = std::max(dmax, nodes[id].depth);
            }
            LevelStat st{dmin, ready.size(), 0, 0.0, 0.0};
            merge_nodes_device(ready, st);
            st.secs = since(tw);
            st.depth_max = dmax;
            st.wave = wave++;
            levels.push_back(st);
            left -= ready.size();
            {
                std::lock_guard<std::mutex> g(avail_mu);
                for (int id : ready) newly_available.push_back(id);
            }
        }
        hosts.join();
        *t_host = std::chrono::duration<double>(t_host_end - t0).count();
        *t_tail = std::chrono::duration<double>(Clock::now() - t_host_end).count();
    }

    void merge_level_host(int depth) {
        std::vector<int> ids;
        for (int i = 0; i < (int)nodes.size(); i++)
            if (!nodes[i].host && nodes[i].depth == depth) ids.push_back(i);
        if (ids.empty()) return;
        auto t0 = Clock::now();
        LevelStat st{depth, ids.size(), 0, 0.0, 0.0};
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (;;) {
                size_t i = next.fetch_add(1);
                if (i >= ids.size()) return;
                Node &n = nodes[ids[i]];
                Node &L = nodes[n.left], &R = nodes[n.right];
                n.P.reset(new Mpz);
                n.Q.reset(new Mpz);
                n.T.reset(new Mpz);
                Mpz y;
                __gmpz_mul(&*n.T, &*L.T, &*R.Q);
                __gmpz_mul(&y, &*L.P, &*R.T);
                __gmpz_add(&*n.T, &*n.T, &y);
                if (ids[i] != 0) __gmpz_mul(&*n.P, &*L.P, &*R.P);
                __gmpz_mul(&*n.Q, &*L.Q, &*R.Q);
                L.P.reset(); L.Q.reset(); L.T.reset();
                R.P.reset(); R.Q.reset(); R.T.reset();
            }
        };
        const int nt = std::max(1, std::min<int>(o.threads, (int)ids.size()));
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; t++) pool.emplace_back(work);
        for (auto &t : pool) t.join();
        st.products = ids.size() * 4 - (ids[0] == 0 ? 1 : 0);
        st.secs = since(t0);
        levels.push_back(st);
    }
};

template <class F>
void parallel_for(size_t n, int threads, F f) {
    std::atomic<size_t> next{0};
    const int workers = (int)std::max<size_t>(1, std::min<size_t>((size_t)threads, n));
    const int share = std::max(1, par::threads_here() / workers);  // helper threads per worker
    auto w = [&] {
        const int saved = par::tl_threads;
        par::tl_threads = share;
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= n) break;
            f(i);
        }
        par::tl_threads = saved;
    };
    std::vector<std::thread> ts;
    for (int t = 1; t < workers; t++) ts.emplace_back(w);
    w();
    for (auto &t : ts) t.join();
}

// Products of the final assembly and the decimal conversion: on the GPU when
// both operands reach min_limbs limbs (libgnt, host buffers), otherwise GMP.
struct Mul {
    gnt_ctx *ctx = nullptr;
    double min_limbs = 1024;        // single products (host buffers, one transfer each way)
    double batch_min_limbs = 1024;  // batched products (packed transfers, shared launches)
    int threads = 1;
    size_t gpu_products = 0, host_products = 0;
    double gpu_limbs = 0;
    double t_gpu = 0;     // seconds inside GPU products (host buffers: copies included)
    double t_gpu_fa = 0;  // of which in the final assembly

    static size_t nl(const gmp_mpz *x) { return (size_t)(x->size < 0 ? -x->size : x->size); }
    bool on_gpu(size_t na, size_t nb) const { return ctx && (double)std::min(na, nb) >= min_limbs; }

    // r = a * b (r may alias a or b; a == b squares)
    void mul(gmp_mpz *r, const gmp_mpz *a, const gmp_mpz *b) {
        const size_t na = nl(a), nb = nl(b);
        if (!on_gpu(na, nb)) {
            __gmpz_mul(r, a, b);
            host_products++;
            return;
        }
        const bool neg = (a->size < 0) != (b->size < 0);
        Mpz t;
        __gmpz_realloc2(&t, (unsigned long)(na + nb) * 64);
        size_t len = 0;
        const auto t0 = Clock::now();
        if (a == b) check(gnt_sqr(ctx, a->d, na, t.z.d, na + nb, &len), "square");
        else check(gnt_mul(ctx, a->d, na, b->d, nb, t.z.d, na + nb, &len), "multiply");
        t_gpu += since(t0);
        t.z.size = neg ? -(int)len : (int)len;
        __gmpz_swap(r, &t);
        gpu_products++;
        gpu_limbs += (double)(na + nb);
    }

    struct Job {
        gmp_mpz *r;
        const gmp_mpz *a, *b;  // nonnegative
    };
    // Independent products: the large ones in one batched GPU call while host
    // threads do the small ones.
    void batch(std::vector<Job> &jobs) {
        std::vector<size_t> g, h;
        for (size_t i = 0; i < jobs.size(); i++)
            (ctx && (double)std::min(nl(jobs[i].a), nl(jobs[i].b)) >= batch_min_limbs ? g : h).push_back(i);
        std::vector<gnt_job> gj(g.size());
        for (size_t k = 0; k < g.size(); k++) {
            Job &j = jobs[g[k]];
            const size_t na = nl(j.a), nb = nl(j.b);
            __gmpz_realloc2(j.r, (unsigned long)(na + nb) * 64);
            gj[k] = gnt_job{j.a->d, na, j.b->d, nb, j.r->d, na + nb, 0, 0, 0};
            gpu_limbs += (double)(na + nb);
        }
        std::thread host([&] { parallel_for(h.size(), std::max(1, threads - 1), [&](size_t k) {
                                   __gmpz_mul(jobs[h[k]].r, jobs[h[k]].a, jobs[h[k]].b);
                               }); });
        const auto t0 = Clock::now();
        if (!gj.empty()) check(gnt_mul_batch(ctx, gj.data(), gj.size()), "batched multiply");
        t_gpu += since(t0);
        host.join();
        for (size_t k = 0; k < g.size(); k++) jobs[g[k]].r->size = (int)gj[k].out_len;
        gpu_products += g.size();
        host_products += h.size();
    }
};

unsigned long bits_of(const gmp_mpz *x) { return x->size ? (unsigned long)__gmpz_sizeinbase(x, 2) : 0; }

// r = x * 2^s (s may be negative: truncation toward zero)
void shift(gmp_mpz *r, const gmp_mpz *x, long s) {
    par::shift(r, x, s);  // multi-threaded on large numbers
}

// Newton precisions from k down to a start the hardware double gives exactly
// enough for (each step at most doubles, minus a margin for rounding).
std::vector<unsigned long> newton_chain(unsigned long k) {
    std::vector<unsigned long> ps;
    for (unsigned long p = k;; p = p / 2 + 6) {
        ps.push_back(p);
        if (p <= 48) break;
    }
    return ps;
}

// R ~ 2^(n+k) / D, n = bits(D): Newton r' = r + r(1 - d r) with precision
// doubling (the reference's reciprocal_at_prec, float.rs:417-474, does the
// same iteration on Floats).  Within a few units of the last place.
void reciprocal(Mul &M, const gmp_mpz *D, unsigned long k, gmp_mpz *R) {
    const unsigned long n = bits_of(D);
    std::vector<unsigned long> ps = newton_chain(k);
    unsigned long p = ps.back();
    {
        Mpz top;
        shift(&top, D, 53 - (long)n);  // 53 leading bits
        const double dt = (double)top.z.d[0];
        const long double r0 = std::ldexp(1.0L / (long double)dt, (int)(p + 53));
        __gmpz_set_ui(R, (unsigned long)r0);
    }
    const unsigned long g = 32;
    Mpz dt, t, e, c;
    for (size_t i = ps.size() - 1; i-- > 0;) {
        const unsigned long q = ps[i];  // next precision
        shift(&dt, D, (long)(q + g) - (long)n);
        M.mul(&t, &dt, R);                              // ~ 2^(q+g+p)
        par::pow2(&e, q + g + p);
        par::sub(&e, &e, &t);                           // ~ err 2^(q+g), signed
        const long eb = (long)bits_of(&e);
        const long drop = std::max(0L, eb - (long)(q - p + 48));
        shift(&e, &e, -drop);
        M.mul(&c, R, &e);
        shift(&c, &c, drop - (long)(2 * p + g));        // r e / 2^(2p+g)
        shift(R, R, (long)(q - p));
        par::add(R, R, &c);
        p = q;
    }
}

// Y ~ 2^k / sqrt(c) for a small integer c: Newton y' = y + y(1 - c y^2)/2.
void inv_sqrt(Mul &M, unsigned long c, unsigned long k, gmp_mpz *Y) {
    std::vector<unsigned long> ps = newton_chain(k);
    unsigned long p = ps.back();
    __gmpz_set_ui(Y, (unsigned long)std::ldexp(1.0L / std::sqrt((long double)c), (int)p));
    Mpz s, e, corr;
    for (size_t i = ps.size() - 1; i-- > 0;) {
        const unsigned long q = ps[i];
        M.mul(&s, Y, Y);                                // ~ 2^(2p) / c
        par::mul_ui(&s, &s, c);
        par::pow2(&e, 2 * p);
        par::sub(&e, &e, &s);                           // ~ err 2^p, signed
        const long eb = (long)bits_of(&e);
        const long drop = std::max(0L, eb - (long)(q - p + 48));
        shift(&e, &e, -drop);
        M.mul(&corr, Y, &e);
        shift(&corr, &corr, drop - (long)(3 * p + 1 - q));
        shift(Y, Y, (long)(q - p));
        par::add(Y, Y, &corr);
        p = q;
    }
}

// pi * 2^prec (within a few units), from the root's Q and |T|:
// pi = 426880 sqrt(10005) Q / (A Q - |T|), every large product through M.
void assemble_pi(Mul &M, const Mpz &Q, const Mpz &Tabs, unsigned long prec, gmp_mpz *pi_fix) {
    const unsigned long w = prec + 64;  // working precision
    Mpz den, q;
    par::mul_ui(&den, &Q, CH_A);
    par::sub(&den, &den, &Tabs);
    __gmpz_set(&q, &Q);
    const long s = (long)bits_of(&den) - (long)(w + 64);
    if (s > 0) {  // Q / den needs w + 64 bits of each
        shift(&den, &den, -s);
        shift(&q, &q, -s);
    }
    const unsigned long n = bits_of(&den), k = w + 32;
    Mpz R, Y, qf;
    reciprocal(M, &den, k, &R);  // ~ 2^(n+k) / den
    inv_sqrt(M, 10005, k, &Y);   // ~ 2^k / sqrt(10005)
    M.mul(&qf, &q, &R);
    shift(&qf, &qf, (long)w - (long)(n + k));  // ~ Q/den 2^w
    par::mul_ui(&Y, &Y, 10005);                // ~ sqrt(10005) 2^k
    M.mul(pi_fix, &Y, &qf);
    par::mul_ui(pi_fix, pi_fix, 426880);
    shift(pi_fix, pi_fix, -(long)(k + w - prec));
}

// Decimal digits of a binary fraction by a scaled remainder tree: the first h
// of n digits of f come from f truncated, the last n - h from the fraction
// part of f * 10^h; only multiplications, a tree level at a time (each level's
// products batched on the GPU).  Every fraction carries an error bound in
// units of its last place; a digit boundary within that bound stops the run
// (the truncation would be undecided), as does the reference's exact
// Round::Down (util.rs:48-56) never need to.
struct Decimal {
    Mul &M;
    int threads;
    size_t leaf_digits = 2048;
    std::map<unsigned long, std::unique_ptr<Mpz>> pow10;

    struct Frac {
        Mpz f;               // value f / 2^F
        unsigned long F = 0;
        size_t n = 0;        // digits wanted
        size_t at = 0;       // offset in the output
        unsigned err = 0;    // error bound, units of 2^-F
    };

    explicit Decimal(Mul &m, int t) : M(m), threads(t) {}

    // 10^h, memoized.  Inserts: call only from one thread at a time.
    const gmp_mpz *p10(unsigned long h) {
        auto it = pow10.find(h);
        if (it != pow10.end()) return &*it->second;
        std::unique_ptr<Mpz> v(new Mpz);
        if (h <= 19) {
            __gmpz_ui_pow_ui(&*v, 10, h);
        } else {
            const gmp_mpz *half = p10(h / 2);
            M.mul(&*v, half, half);
            if (h & 1) __gmpz_mul_ui(&*v, &*v, 10);
        }
        const gmp_mpz *r = &*v;
        pow10[h] = std::move(v);
        return r;
    }

    // x in [E 10^m, 2^F - E 10^m]: the digits cannot move under the error E.
    bool decided(const gmp_mpz *rem, unsigned long F, unsigned err, const gmp_mpz *p10m) {
        Mpz b, top;
        par::mul_ui(&b, p10m, err + 1);
        if (__gmpz_cmp(rem, &b) < 0) return false;
        par::pow2(&top, F);
        par::sub(&top, &top, &b);
        return __gmpz_cmp(rem, &top) <= 0;
    }

    void leaf(Frac &x, std::string &out) {
        Mpz v, digits, rem;
        const gmp_mpz *pw = p10(x.n);
        __gmpz_mul(&v, &x.f, pw);
        __gmpz_tdiv_q_2exp(&digits, &v, x.F);
        __gmpz_tdiv_r_2exp(&rem, &v, x.F);
        if (!decided(&rem, x.F, x.err, pw)) die("decimal conversion undecided at a leaf", "");
        char *s = __gmpz_get_str(nullptr, 10, &digits);
        const size_t len = std::strlen(s);
        if (len > x.n) die("leaf overflow", "");
        std::memset(&out[x.at], '0', x.n - len);
        std::memcpy(&out[x.at + x.n - len], s, len);
        std::free(s);
    }

    // n digits of f / 2^F into out[0..n)
    std::string stats;  // per level: nodes, products on the GPU, seconds (JSON)

    void run(Frac root, std::string &out) {
        std::vector<Frac> level;
        level.push_back(std::move(root));
        std::vector<Frac> leaves;
        stats = "[";
        while (!level.empty()) {
            auto t0 = Clock::now();
            const size_t gpu0 = M.gpu_products;
            std::vector<Frac> next;
            std::vector<size_t> split;
            for (size_t i = 0; i < level.size(); i++)
                if (level[i].n <= leaf_digits) leaves.push_back(std::move(level[i]));
                else split.push_back(i);
            if (split.empty()) break;
            // powers first (sequential: they share squares), then the level's products
            for (size_t i : split) p10(level[i].n / 2);
            const double t_pow = since(t0);
            std::vector<Mpz> g(split.size());
            std::vector<Mul::Job> jobs;
            for (size_t k = 0; k < split.size(); k++)
                jobs.push_back(Mul::Job{&g[k], &level[split[k]].f, p10(level[split[k]].n / 2)});
            M.batch(jobs);
            const double t_mul = since(t0) - t_pow;
            next.resize(2 * split.size());
            parallel_for(split.size(), threads, [&](size_t k) {
                Frac &x = level[split[k]];
                const size_t h = x.n / 2, l = x.n - h;
                const unsigned long bh = (unsigned long)std::ceil((double)h * std::log2(10.0));
                const unsigned long bl = (unsigned long)std::ceil((double)l * std::log2(10.0));
                Frac &hi = next[2 * k], &lo = next[2 * k + 1];
                // high digits: f truncated by the bits the low digits need
                hi.F = x.F - bl;
                shift(&hi.f, &x.f, -(long)bl);
                hi.n = h;
                hi.at = x.at;
                hi.err = x.err + 1;
                // low digits: fraction part of f 10^h
                Mpz rem;
                par::trunc(&rem, &g[k], x.F);
                if (!decided(&rem, x.F, x.err, p10(h))) die("decimal conversion undecided", "");
                lo.F = x.F - bh;
                shift(&lo.f, &rem, -(long)bh);
                lo.n = l;
                lo.at = x.at + h;
                lo.err = x.err + 1;
            });
            char buf[200];
            std::snprintf(buf, sizeof buf, "%s{\"nodes\": %zu, \"gpu_products\": %zu, \"s_powers\": %.4f, \"s_products\": %.4f, \"s\": %.4f}",
                          stats.size() > 1 ? ", " : "", split.size(), M.gpu_products - gpu0, t_pow, t_mul, since(t0));
            stats += buf;
            level = std::move(next);
        }
        // every power the leaves need exists before the threads start: p10()
        // inserts into the map, which is not safe to do concurrently
        auto t0 = Clock::now();
        for (const Frac &x : leaves) p10(x.n);
        parallel_for(leaves.size(), threads, [&](size_t i) { leaf(leaves[i], out); });
        char buf[120];
        std::snprintf(buf, sizeof buf, "%s{\"leaves\": %zu, \"s\": %.4f}]", stats.size() > 1 ? ", " : "", leaves.size(),
                      since(t0));
        stats += buf;
    }
};

// floor(pi * 10^(D-1)) as D characters with every large product on the GPU.
std::string g_decimal_stats = "[]";  // last conversion's per-level timings (JSON)

// D characters of pi from pi_fix ~ pi 2^prec (error bound err units).
std::string digits_of(Mul &M, const Mpz &pi_fix, uint64_t digits, uint64_t prec, int threads, unsigned err) {
    Mpz ip;
    par::shift(&ip, &pi_fix, -(long)prec);
    if (ip.limbs() != 1 || ip.z.d[0] != 3) die("integer part is not 3", "");
    std::string out(digits, '0');
    out[0] = '3';
    if (digits > 1) {
        Decimal dec(M, threads);
        Decimal::Frac root;
        par::trunc(&root.f, &pi_fix, prec);
        root.F = prec;
        root.n = digits - 1;
        root.at = 1;
        root.err = err;
        dec.run(std::move(root), out);
        g_decimal_stats = dec.stats;
    }
    return out;
}

std::string pi_digits_gpu(Mul &M, const Mpz &Q, const Mpz &Tabs, uint64_t digits, uint64_t prec, int threads,
                          double *t_fa, double *t_dec) {
    auto t0 = Clock::now();
    Mpz pi_fix;
    assemble_pi(M, Q, Tabs, (unsigned long)prec, &pi_fix);
    *t_fa = since(t0);
    M.t_gpu_fa = M.t_gpu;
    t0 = Clock::now();
    std::string out = digits_of(M, pi_fix, digits, prec, threads, 4);
    *t_dec = since(t0);
    return out;
}

// ---- Gauss-Legendre (gauss_legendre.rs:61-166) ----

// ~ sqrt(x), within a small fraction of a unit: Newton for y ~ 1/sqrt(v),
// v = x / 2^e in [1/4, 1) (e even), then sqrt(x) = x y 2^(-e/2).  The
// reference's Float::sqrt_mut (float.rs:85-174) is the same precision-
// doubling iteration on Floats.
void sqrt_big(Mul &M, const gmp_mpz *x, unsigned long k, gmp_mpz *out) {
    const unsigned long n = bits_of(x);
    const unsigned long e = n + (n & 1);
    std::vector<unsigned long> ps = newton_chain(k);
    unsigned long p = ps.back();
    Mpz y;
    {
        Mpz top;
        shift(&top, x, 53 - (long)n);
        const long double v = std::ldexp((long double)top.z.d[0], (int)n - 53 - (int)e);  // x / 2^e
        __gmpz_set_ui(&y, (unsigned long)std::ldexp(1.0L / std::sqrt(v), (int)p));
    }
    const unsigned long g = 32;
    Mpz vt, sq, t, er, corr;
    for (size_t i = ps.size() - 1; i-- > 0;) {
        const unsigned long q = ps[i];
        shift(&vt, x, (long)(q + g) - (long)e);   // ~ v 2^(q+g)
        M.mul(&sq, &y, &y);                        // ~ 2^(2p) / v
        M.mul(&t, &vt, &sq);                       // ~ 2^(q+g+2p)
        par::pow2(&er, q + g + 2 * p);
        par::sub(&er, &er, &t);                    // ~ err 2^(q+g+p), signed
        const long eb = (long)bits_of(&er);
        const long drop = std::max(0L, eb - (long)(q - p + 48));
        shift(&er, &er, -drop);
        M.mul(&corr, &y, &er);
        shift(&corr, &corr, drop - (long)(g + 3 * p + 1));
        shift(&y, &y, (long)(q - p));
        par::add(&y, &y, &corr);
        p = q;
    }
    // sqrt(x) = x y_k / 2^(e/2 + k), with x cut to k + 64 bits
    Mpz xt;
    const long d = std::max(0L, (long)n - (long)(k + 64));
    shift(&xt, x, -d);
    M.mul(out, &xt, &y);
    shift(out, out, d - (long)(e / 2 + k));
}

// pi 2^prec by the Brent-Salamin AGM in fixed point with 64 guard bits:
//   a0 = 1, b0 = 1/sqrt(2), t0 = 1/4, p0 = 1
//   a' = (a+b)/2, b' = sqrt(a b), t' = t - p (a - a')^2, p' = 2p
//   pi = (a+b)^2 / (4t)
// for ceil(log2 prec) + 2 iterations (gauss_legendre.rs:77-79, 119-153).
void pi_gauss_legendre(Mul &M, unsigned long prec, gmp_mpz *pi_fix, int *iters_out) {
    const unsigned long W = prec + 64, k = W + 32;
    Mpz a, b, t, an, ab, dd;
    __gmpz_set_ui(&a, 1);
    __gmpz_mul_2exp(&a, &a, W);
    inv_sqrt(M, 2, W, &b);                 // 2^W / sqrt(2)
    __gmpz_set_ui(&t, 1);
    __gmpz_mul_2exp(&t, &t, W - 2);
    const int iters = (int)std::ceil(std::log2((double)prec)) + 2;
    for (int i = 0; i < iters; i++) {
        par::add(&an, &a, &b);
        shift(&an, &an, -1);
        M.mul(&ab, &a, &b);                // scale 2^(2W)
        sqrt_big(M, &ab, k, &b);           // b' = sqrt(a b), scale 2^W
        par::sub(&dd, &a, &an);
        M.mul(&dd, &dd, &dd);
        shift(&dd, &dd, (long)i - (long)W);  // p (a - a')^2 with p = 2^i
        par::sub(&t, &t, &dd);
        __gmpz_swap(&a, &an);
    }
    par::add(&an, &a, &b);
    M.mul(&ab, &an, &an);                  // (a+b)^2, scale 2^(2W)
    shift(&ab, &ab, -(long)W);
    Mpz R;
    const unsigned long nt = bits_of(&t);
    reciprocal(M, &t, k, &R);              // ~ 2^(nt+k) / t
    M.mul(pi_fix, &ab, &R);                // (a+b)^2 / (4t) at scale 2^W
    shift(pi_fix, pi_fix, (long)W - (long)(nt + k + 2) - 64);
    *iters_out = iters;
}

// floor(pi * 10^(D-1)) with GMP's own square root, division and radix
// conversion (comparison path).
std::string final_assembly_gmp(const Mpz &Q, const Mpz &Tabs, uint64_t digits, uint64_t prec, double *t_fa,
                           double *t_dec) {
    auto t0 = Clock::now();
    // denom = A Q + T = A Q - |T|
    Mpz den;
    __gmpz_mul_ui(&den, &Q, CH_A);
    __gmpz_sub(&den, &den, &Tabs);
    // Q / den needs only prec + 64 bits of each
    Mpz q;
    __gmpz_set(&q, &Q);
    const size_t dbits = __gmpz_sizeinbase(&den, 2);
    if (dbits > prec + 64) {
        const unsigned long s = (unsigned long)(dbits - (prec + 64));
        __gmpz_tdiv_q_2exp(&q, &q, s);
        __gmpz_tdiv_q_2exp(&den, &den, s);
    }
    // S = floor(sqrt(10005) 2^prec)
    Mpz S;
    __gmpz_set_ui(&S, 10005);
    __gmpz_mul_2exp(&S, &S, 2 * prec);
    __gmpz_sqrt(&S, &S);
    // pi_fix = floor(426880 S q / den), within 2 units below pi 2^prec
    Mpz pifix;
    __gmpz_mul(&pifix, &S, &q);
    __gmpz_mul_ui(&pifix, &pifix, 426880);
    __gmpz_tdiv_q(&pifix, &pifix, &den);
    *t_fa = since(t0);
    t0 = Clock::now();
    // floor(pi 10^(D-1)): scale, then drop prec bits; the dropped bits must
    // stay 2*10^(D-1) away from the next integer (the error bound above)
    Mpz p10, scaled, intpart, frac, bound, top;
    __gmpz_ui_pow_ui(&p10, 10, (unsigned long)(digits - 1));
    __gmpz_mul(&scaled, &pifix, &p10);
    __gmpz_tdiv_q_2exp(&intpart, &scaled, prec);
    __gmpz_tdiv_r_2exp(&frac, &scaled, prec);
    __gmpz_mul_ui(&bound, &p10, 2);
    __gmpz_set_ui(&top, 1);
    __gmpz_mul_2exp(&top, &top, prec);
    __gmpz_sub(&top, &top, &bound);
    if (__gmpz_cmp(&frac, &top) >= 0) die("truncation undecided at this precision", "");
    char *s = __gmpz_get_str(nullptr, 10, &intpart);
    std::string out(s);
    std::free(s);
    *t_dec = since(t0);
    if (out.size() != digits) die("unexpected digit count", "");
    return out;
}

void usage() {
    std::fprintf(stderr,
                 "usage: pi_compute DIGITS OUT_FILE [--algorithm chudnovsky|gauss-legendre] [--gpu-min-limbs L]\n"
                 "                  [--threads T] [--host-merge] [--merge dataflow|levels] [--shutdown]\n"
                 "                     [--fa-gmp] [--fa-min-limbs L] [--dec-min-limbs L] [--par-min-limbs L] [--device D]\n"
                 "                     [--plans CACHE.json]\n");
    std::exit(2);
}

}  // namespace

// End of a run: output flushed; the runtime's exit releases what the
// context still holds.
int finish(const Options &) {
    std::fflush(stdout);
    std::fflush(stderr);
    return 0;
}

void write_digits(FILE *f, const std::string &s, uint64_t digits, const std::string &path) {
    std::fwrite(s.data(), 1, 1, f);
    if (digits > 1) {
        std::fputc('.', f);
        std::fwrite(s.data() + 1, 1, s.size() - 1, f);
    }
    std::fputc('\n', f);
    if (std::fclose(f) != 0) die("write failed", path.c_str());
}

int run_gauss_legendre(const Options &o, uint64_t prec, FILE *f, Clock::time_point t_start) {
    gnt_ctx *ctx = nullptr;
    auto t0 = Clock::now();
    const bool gpu = (double)prec / 64.0 >= o.fa_min_limbs;
    if (gpu) {
        gnt_config cfg{};
        cfg.device = o.device;
        cfg.plan_cache_path = o.plans.empty() ? nullptr : o.plans.c_str();
        check(gnt_init(&ctx, &cfg), "init");
    }
    const double t_init = since(t0);
    Mul M;
    M.ctx = ctx;
    M.min_limbs = o.fa_min_limbs;
    M.batch_min_limbs = std::max(o.dec_min_limbs, 1.0);
    M.threads = o.threads;
    t0 = Clock::now();
    Mpz pi_fix;
    int iters = 0;
    pi_gauss_legendre(M, (unsigned long)prec, &pi_fix, &iters);
    const double t_agm = since(t0);
    t0 = Clock::now();
    // AGM rounding: t loses up to ~2^iters units at 2^-W, W = prec + 64
    const std::string s = digits_of(M, pi_fix, o.digits, prec, o.threads, 8);
    const double t_dec = since(t0);
    t0 = Clock::now();
    write_digits(f, s, o.digits, o.out);
    const double t_write = since(t0);
    if (ctx && o.shutdown) gnt_shutdown(ctx);  // else the process exit returns the device memory
    else if (ctx) gnt_sync(ctx);
    std::printf("{\"digits\": %llu, \"algorithm\": \"gauss-legendre\", \"prec_bits\": %llu, \"iterations\": %d, "
                "\"threads\": %d, \"products\": \"%s\", \"t_init\": %.3f, \"t_agm\": %.3f, \"t_decimal\": %.3f, "
                "\"t_write\": %.3f, \"t_total\": %.3f, \"gpu_products\": %zu, \"host_products\": %zu, "
                "\"gpu_operand_limbs\": %.0f}\n",
                (unsigned long long)o.digits, (unsigned long long)prec, iters, o.threads, gpu ? "gpu" : "host", t_init,
                t_agm, t_dec, t_write, since(t_start), M.gpu_products, M.host_products, M.gpu_limbs);
    return finish(o);
}

int main(int argc, char **argv) {
    if (argc < 3) usage();
    Options o;
    o.digits = std::strtoull(argv[1], nullptr, 10);
    o.out = argv[2];
    for (int i = 3; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--gpu-min-limbs" && i + 1 < argc) o.gpu_min_limbs = std::atof(argv[++i]);
        else if (a == "--threads" && i + 1 < argc) o.threads = std::atoi(argv[++i]);
        else if (a == "--host-merge") o.host_merge = true;
        else if (a == "--fa-gmp") o.fa_gmp = true;
        else if (a == "--fa-min-limbs" && i + 1 < argc) o.fa_min_limbs = std::atof(argv[++i]);
        else if (a == "--dec-min-limbs" && i + 1 < argc) o.dec_min_limbs = std::atof(argv[++i]);
        else if (a == "--device" && i + 1 < argc) o.device = std::atoi(argv[++i]);
        else if (a == "--plans" && i + 1 < argc) o.plans = argv[++i];
        else if (a == "--algorithm" && i + 1 < argc) o.algorithm = argv[++i];
        else if (a == "--shutdown") o.shutdown = true;
        else if (a == "--par-min-limbs" && i + 1 < argc) o.par_min_limbs = std::atof(argv[++i]);
        else if (a == "--merge" && i + 1 < argc) {
            const std::string m = argv[++i];
            if (m == "levels") o.merge_levels = true;
            else if (m != "dataflow") usage();
        }
        else usage();
    }
    if (o.threads <= 0) {
        unsigned hc = std::thread::hardware_concurrency();
        o.threads = (int)std::min(16u, hc ? hc : 1u);
    }
    par::cfg.threads = o.threads;
    par::cfg.min_limbs = (size_t)std::max(1.0, o.par_min_limbs);
    auto t_start = Clock::now();
    FILE *f = std::fopen(o.out.c_str(), "wb");
    if (!f) die("cannot open", o.out.c_str());
    if (o.digits == 0) {  // chudnovsky.rs:82-85: only the sink's finish
        std::fputc('\n', f);
        std::fclose(f);
        std::printf("{\"digits\": 0}\n");
        return 0;
    }
    const uint64_t n_terms = (uint64_t)std::ceil((double)o.digits / DIGITS_PER_TERM) + SAFETY_TERMS;
    const uint64_t prec = (uint64_t)std::ceil((double)o.digits * std::log2(10.0)) + SAFETY_BITS;
    if (o.algorithm == "gauss-legendre") return run_gauss_legendre(o, prec, f, t_start);
    if (o.algorithm != "chudnovsky") usage();

    Driver d;
    d.o = o;
    d.nodes.reserve(1024);
    d.build(1, n_terms + 1, 0);
    const bool any_merge = !d.nodes[0].host;

    double t_init = 0, t_upload = 0, t_download = 0;
    const bool gpu_merge = any_merge && d.merges_on_device();
    const bool gpu_fa = !o.fa_gmp && (double)prec / 64.0 >= o.fa_min_limbs;
    if (gpu_merge || gpu_fa) {
        auto t0 = Clock::now();
        gnt_config cfg{};
        cfg.device = o.device;
        cfg.plan_cache_path = o.plans.empty() ? nullptr : o.plans.c_str();
        // the merges want throughput, not the tail latency that one large
        // product at a time buys a stream of independent requests
        cfg.serve_large_inflight = SIZE_MAX;
        check(gnt_init(&d.ctx, &cfg), "init");
        if (gpu_merge) check(gnt_serve_start(d.ctx), "serve start");
        t_init = since(t0);
    }
    auto t0 = Clock::now();
    double t_host = 0, t_merge = 0;
    const bool dataflow = gpu_merge && !o.merge_levels;
    if (dataflow) {
        // merges as children finish; t_merge is what is left after the host nodes
        d.run_dataflow(o.threads, &t_host, &t_merge);
    } else {
        d.run_host_nodes(o.threads, gpu_merge);  // merged on the device: uploads overlap the host work
        t_host = since(t0);
    }

    Mpz Q, Tabs;
    if (!any_merge) {
        __gmpz_set(&Q, &*d.nodes[0].Q);
        __gmpz_neg(&Tabs, &*d.nodes[0].T);
    } else if (!d.merges_on_device()) {
        t0 = Clock::now();
        for (int depth = d.max_depth; depth >= 0; depth--) d.merge_level_host(depth);
        t_merge = since(t0);
        __gmpz_set(&Q, &*d.nodes[0].Q);
        __gmpz_neg(&Tabs, &*d.nodes[0].T);
    } else {
        if (!dataflow) {
            t0 = Clock::now();
            for (int depth = d.max_depth; depth >= 0; depth--) d.merge_level_device(depth);
            t_merge = since(t0);
        }
        t0 = Clock::now();
        Node &root = d.nodes[0];
        size_t nq = Driver::buf_len(d.ctx, root.bq), nt = Driver::buf_len(d.ctx, root.bt);
        size_t got = 0;
        __gmpz_realloc2(&Q, (unsigned long)(nq ? nq : 1) * 64);
        __gmpz_realloc2(&Tabs, (unsigned long)(nt ? nt : 1) * 64);
        check(gnt_buf_download(d.ctx, root.bq, Q.z.d, nq, &got), "download");
        Q.z.size = (int)got;
        check(gnt_buf_download(d.ctx, root.bt, Tabs.z.d, nt, &got), "download");
        Tabs.z.size = (int)got;
        gnt_buf_release(d.ctx, root.bq);
        gnt_buf_release(d.ctx, root.bt);
        t_download = since(t0);
    }

    uint64_t sstats[6] = {0, 0, 0, 0, 0, 0};
    double t_serve_stop = 0;
    if (gpu_merge) {
        t0 = Clock::now();
        gnt_serve_stats(d.ctx, sstats, 6);
        gnt_serve_stop(d.ctx);
        t_serve_stop = since(t0);
    }
    double t_fa = 0, t_dec = 0;
    Mul M;
    M.ctx = gpu_fa ? d.ctx : nullptr;
    M.min_limbs = o.fa_min_limbs;
    M.batch_min_limbs = std::max(o.dec_min_limbs, 1.0);
    M.threads = o.threads;
    std::string s = o.fa_gmp ? final_assembly_gmp(Q, Tabs, o.digits, prec, &t_fa, &t_dec)
                             : pi_digits_gpu(M, Q, Tabs, o.digits, prec, o.threads, &t_fa, &t_dec);
    t0 = Clock::now();
    write_digits(f, s, o.digits, o.out);
    const double t_write = since(t0);
    // The process ends here and the runtime's exit releases the device
    // memory; gnt_shutdown would first free the pools buffer by buffer
    // (~0.3 s more at 10^8 digits).  Only the stream is drained.
    t0 = Clock::now();
    if (d.ctx && o.shutdown) gnt_shutdown(d.ctx);
    else if (d.ctx) gnt_sync(d.ctx);
    const double t_shutdown = since(t0);
    const double t_total = since(t_start);
    double t_levels = 0;
    std::string lv = "[";
    for (size_t i = 0; i < d.levels.size(); i++) {
        const LevelStat &L = d.levels[i];
        t_levels += L.secs;
        char buf[256];
        if (L.wave >= 0)  // a merge wave: every node whose children were ready, of depths [depth, depth_max]
            std::snprintf(buf, sizeof buf,
                          "%s{\"wave\": %d, \"depth\": %d, \"depth_max\": %d, \"nodes\": %zu, \"products\": %zu, "
                          "\"operand_limbs\": %.0f, \"s\": %.4f}",
                          i ? ", " : "", L.wave, L.depth, L.depth_max, L.nodes, L.products, L.limbs, L.secs);
        else
            std::snprintf(buf, sizeof buf, "%s{\"depth\": %d, \"nodes\": %zu, \"products\": %zu, \"operand_limbs\": %.0f, \"s\": %.4f}",
                          i ? ", " : "", L.depth, L.nodes, L.products, L.limbs, L.secs);
        lv += buf;
    }
    lv += "]";
    std::printf("{\"digits\": %llu, \"n_terms\": %llu, \"prec_bits\": %llu, \"merge\": \"%s\", \"merge_order\": \"%s\", "
                "\"merge_busy_s\": %.3f, \"threads\": %d, "
                "\"gpu_min_limbs\": %.0f, \"fa_min_limbs\": %.0f, \"dec_min_limbs\": %.0f, \"host_nodes\": %zu, \"merge_nodes\": %zu, \"t_init\": %.3f, "
                "\"t_host_bs\": %.3f, \"t_upload\": %.3f, \"t_merge\": %.3f, \"t_download\": %.3f, "
                "\"t_serve_stop\": %.3f, \"t_final_assembly\": %.3f, \"t_decimal\": %.3f, \"t_write\": %.3f, "
                "\"t_shutdown\": %.3f, \"t_total\": %.3f, "
                "\"serve_iterations\": %llu, \"serve_launches\": %llu, \"fa\": \"%s\", \"fa_gpu_products\": %zu, "
                "\"fa_host_products\": %zu, \"fa_gpu_operand_limbs\": %.0f, \"fa_gpu_s\": %.3f, \"dec_gpu_s\": %.3f, "
                "\"decimal_levels\": %s, "
                "\"levels\": %s}\n",
                (unsigned long long)o.digits, (unsigned long long)n_terms, (unsigned long long)prec,
                !any_merge ? "none" : (d.merges_on_device() ? "gpu" : "host"),
                dataflow ? "dataflow" : "levels", t_levels, o.threads, o.gpu_min_limbs,
                o.fa_min_limbs, o.dec_min_limbs, d.host_nodes.size(), d.nodes.size() - d.host_nodes.size(), t_init, t_host, t_upload, t_merge,
                t_download, t_serve_stop, t_fa, t_dec, t_write, t_shutdown, t_total, (unsigned long long)sstats[0],
                (unsigned long long)sstats[1], o.fa_gmp ? "gmp" : (gpu_fa ? "gpu" : "host"), M.gpu_products,
                M.host_products, M.gpu_limbs, M.t_gpu_fa, M.t_gpu - M.t_gpu_fa, g_decimal_stats.c_str(), lv.c_str());
    return finish(o);
}
