// sfem_gnn_host.cpp -- the learned block preconditioner: network parameters,
// forward / backward orchestration, residual-solution pair collection with the
// solver, Adam training, and the C-ABI entry points (sfem_gnn_*).
//
// Network (per node, width 32, 3 message-passing rounds):
//   x (16 features) -> h0 = W_e2 relu(W_e1 x + b) + b
//   round r: m = mean_{j in N(i)} h_j ; h += W2 relu(W1 [h | m] + b1) + b2
//   o = W_d2 relu(W_d1 h + b) + b  (6 used of 16)
//   L = lower(exp o0..2, o3..5), Minv = C (L L^T + eps I) C^T, C C^T = D^-1
// with D the block-Jacobi block.  The last layer starts at zero, so the
// untrained network is the block-Jacobi preconditioner (scaled by 1 + eps).
//
// Loss over residual/solution pairs (r, e) with K e = r collected by the
// solver: one optimally scaled preconditioned step from zero leaves the
// energy-norm error fraction
//   l = 1 - (z.r)^2 / ((z.K z)(r.e)),   z = Minv r,
// averaged over a minibatch.  K z is the device Hessian-vector product.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "sfem_gnn.hpp"
#include "sfem_internal.hpp"

using namespace sfem;

namespace {

constexpr int WIDTH = 32, FIN = 16, FOUT = 16, ROUNDS = 3;
constexpr float BLOCK_EPS = 0.01f;
constexpr int NPARTS = 256;      // workgroups of the weight-gradient reductions
constexpr int MAX_BATCH = 16;

struct Layer {
    int K, NO;
    size_t w, b;      // offsets into the FP32 parameter vector
    size_t wb;        // offset into the bf16 copies (W and W^T)
};

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

struct sfem_gnn {
    std::vector<Layer> L;   // e1, e2, (r1, r2) x ROUNDS, d1, d2
    size_t nparam = 0, nbf = 0;
    std::vector<float> host_w;
    DevBuf w, g, m, v, wb, wtb;
    int t = 0;   // Adam step count
    hipStream_t stream = nullptr;

    // workspaces (sized for the largest node count seen)
    int64_t ws_n = 0;
    DevBuf feat, cbase, A1, H1, h[ROUNDS + 1], Mg[ROUNDS], B[ROUNDS], C[ROUNDS], A3, H3, O, minv;
    DevBuf dO, dT1, dT2, dh, dU, dm, dMinv, part;
    int ws_s = 0;
    DevBuf rb, zb, qb, dpart, c12;

    ~sfem_gnn()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }

    const Layer& e1() const { return L[0]; }
    const Layer& e2() const { return L[1]; }
    const Layer& r1(int r) const { return L[2 + 2 * r]; }
    const Layer& r2(int r) const { return L[3 + 2 * r]; }
    const Layer& d1() const { return L[2 + 2 * ROUNDS]; }
    const Layer& d2() const { return L[3 + 2 * ROUNDS]; }

    float* W(const Layer& l) const { return w.as<float>() + l.w; }
    float* Bias(const Layer& l) const { return w.as<float>() + l.b; }
    float* gW(const Layer& l) const { return g.as<float>() + l.w; }
    float* gB(const Layer& l) const { return g.as<float>() + l.b; }
    const uint16_t* Wb(const Layer& l) const { return wb.as<uint16_t>() + l.wb; }
    const uint16_t* WTb(const Layer& l) const { return wtb.as<uint16_t>() + l.wb; }

    void init(uint64_t seed)
    {
        const int dims[][2] = {{FIN, WIDTH}, {WIDTH, WIDTH}};
        auto add = [&](int K, int NO) {
            Layer l{K, NO, nparam, nparam + (size_t)K * NO, nbf};
            nparam += (size_t)K * NO + NO;
            nbf += (size_t)K * NO;
            L.push_back(l);
        };
        for (auto& d : dims) add(d[0], d[1]);
        for (int r = 0; r < ROUNDS; r++) {
            add(2 * WIDTH, WIDTH);
            add(WIDTH, WIDTH);
        }
        add(WIDTH, WIDTH);
        add(WIDTH, FOUT);
        host_w.assign(nparam, 0.f);
        uint64_t s = seed;
        for (size_t li = 0; li + 1 < L.size(); li++) {   // He-uniform; last layer stays zero
            const Layer& l = L[li];
            const double a = std::sqrt(6.0 / l.K);
            for (int k = 0; k < l.K * l.NO; k++) {
                const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
                host_w[l.w + k] = (float)((2 * u - 1) * a);
            }
        }
        HIPCK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        w.upload(host_w);
        g.alloc(nparam * sizeof(float));
        m.alloc(nparam * sizeof(float));
        v.alloc(nparam * sizeof(float));
        HIPCK(hipMemset(g.p, 0, nparam * sizeof(float)));
        HIPCK(hipMemset(m.p, 0, nparam * sizeof(float)));
        HIPCK(hipMemset(v.p, 0, nparam * sizeof(float)));
        wb.alloc(nbf * sizeof(uint16_t));
        wtb.alloc(nbf * sizeof(uint16_t));
        refresh_bf16();
        HIPCK(hipStreamSynchronize(stream));
    }

    void refresh_bf16()
    {
        for (const Layer& l : L)
            HIPCK(gnn_to_bf16_pair(l.K, l.NO, W(l), wb.as<uint16_t>() + l.wb, wtb.as<uint16_t>() + l.wb, stream));
    }

    void workspace(int64_t n)
    {
        if (n <= ws_n) return;
        ws_n = n;
        const size_t f = sizeof(float);
        feat.alloc(n * FIN * f);
        cbase.alloc(n * 6 * f);
        A1.alloc(n * WIDTH * f);
        H1.alloc(n * WIDTH * f);
        for (auto& b : h) b.alloc(n * WIDTH * f);
        for (int r = 0; r < ROUNDS; r++) {
            Mg[r].alloc(n * WIDTH * f);
            B[r].alloc(n * WIDTH * f);
            C[r].alloc(n * WIDTH * f);
        }
        A3.alloc(n * WIDTH * f);
        H3.alloc(n * WIDTH * f);
        O.alloc(n * FOUT * f);
        minv.alloc(n * 9 * f);
        dO.alloc(n * FOUT * f);
        dT1.alloc(n * WIDTH * f);
        dT2.alloc(n * WIDTH * f);
        dh.alloc(n * WIDTH * f);
        dU.alloc(n * 2 * WIDTH * f);
        dm.alloc(n * WIDTH * f);
        dMinv.alloc(n * 9 * f);
        part.alloc((size_t)NPARTS * 2 * WIDTH * WIDTH * f);
    }

    void batch_space(sfem_ctx* c, int S)
    {
        workspace(c->nn);
        const size_t n3 = 3 * (size_t)c->nn;
        if (rb.n < S * n3 * sizeof(float) || S > ws_s) {
            rb.alloc(S * n3 * sizeof(float));
            zb.alloc(S * n3 * sizeof(float));
            qb.alloc(S * n3 * sizeof(float));
            ws_s = S;
        }
        if (!dpart.p) {
            dpart.alloc((size_t)NPARTS * MAX_BATCH * 2 * sizeof(double));
            c12.alloc(2 * MAX_BATCH * sizeof(float));
        }
    }

    // features of the context (Jacobi blocks in c->d_bd32 must be current)
    // train: also keep the pre-activations the backward pass reads
    void forward(sfem_ctx* c, hipStream_t s, bool train = true)
    {
        const int64_t N = c->nn;
        workspace(N);
        const int* ap = c->d_adj_ptr.as<int>();
        const int* ai = c->d_adj_idx.as<int>();
        HIPCK(gnn_features(N, c->d_bd32.as<float>(), c->d_fixed.as<uint8_t>(), ap, ai, feat.as<float>(), cbase.as<float>(), s));
        HIPCK(gnn_gemm(FIN, WIDTH, N, feat.as<float>(), FIN, nullptr, Wb(e1()), Bias(e1()), train ? A1.as<float>() : nullptr,
                       H1.as<float>(), s));
        HIPCK(gnn_gemm(WIDTH, WIDTH, N, H1.as<float>(), WIDTH, nullptr, Wb(e2()), Bias(e2()), h[0].as<float>(), nullptr, s));
        for (int r = 0; r < ROUNDS; r++) {
            HIPCK(gnn_mean_fwd(N, ap, ai, h[r].as<float>(), Mg[r].as<float>(), s));
            HIPCK(gnn_gemm(2 * WIDTH, WIDTH, N, h[r].as<float>(), WIDTH, Mg[r].as<float>(), Wb(r1(r)), Bias(r1(r)),
                           train ? B[r].as<float>() : nullptr, C[r].as<float>(), s));
            // h_{r+1} = h_r + (C_r W + b): the residual add in the product's epilogue
            HIPCK(gnn_gemm(WIDTH, WIDTH, N, C[r].as<float>(), WIDTH, nullptr, Wb(r2(r)), Bias(r2(r)), h[r + 1].as<float>(),
                           nullptr, s, h[r].as<float>()));
        }
        HIPCK(gnn_gemm(WIDTH, WIDTH, N, h[ROUNDS].as<float>(), WIDTH, nullptr, Wb(d1()), Bias(d1()),
                       train ? A3.as<float>() : nullptr, H3.as<float>(), s));
        HIPCK(gnn_gemm(WIDTH, FOUT, N, H3.as<float>(), WIDTH, nullptr, Wb(d2()), Bias(d2()), O.as<float>(), nullptr, s));
        HIPCK(gnn_blocks_fwd(N, O.as<float>(), cbase.as<float>(), BLOCK_EPS, minv.as<float>(), s));
    }

    // parameter gradient from dMinv (forward must have run)
    void backward(sfem_ctx* c, hipStream_t s)
    {
        const int64_t N = c->nn;
        const int* ap = c->d_adj_ptr.as<int>();
        const int* ai = c->d_adj_idx.as<int>();
        float* P = part.as<float>();
        HIPCK(gnn_blocks_bwd(N, O.as<float>(), cbase.as<float>(), dMinv.as<float>(), dO.as<float>(), s));
        // decoder
        HIPCK(gnn_wgrad(WIDTH, FOUT, N, H3.as<float>(), WIDTH, nullptr, dO.as<float>(), P, NPARTS, gW(d2()), s));
        HIPCK(gnn_bgrad(FOUT, N, dO.as<float>(), P, NPARTS, gB(d2()), s));
        HIPCK(gnn_gemm(FOUT, WIDTH, N, dO.as<float>(), FOUT, nullptr, WTb(d2()), nullptr, dT1.as<float>(), nullptr, s));
        HIPCK(gnn_relu_bwd(N * WIDTH, A3.as<float>(), dT1.as<float>(), dT2.as<float>(), s));
        HIPCK(gnn_wgrad(WIDTH, WIDTH, N, h[ROUNDS].as<float>(), WIDTH, nullptr, dT2.as<float>(), P, NPARTS, gW(d1()), s));
        HIPCK(gnn_bgrad(WIDTH, N, dT2.as<float>(), P, NPARTS, gB(d1()), s));
        HIPCK(gnn_gemm(WIDTH, WIDTH, N, dT2.as<float>(), WIDTH, nullptr, WTb(d1()), nullptr, dh.as<float>(), nullptr, s));
        // message-passing rounds, last to first; dh holds dL/dh_{r+1}
        for (int r = ROUNDS - 1; r >= 0; r--) {
            HIPCK(gnn_wgrad(WIDTH, WIDTH, N, C[r].as<float>(), WIDTH, nullptr, dh.as<float>(), P, NPARTS, gW(r2(r)), s));
            HIPCK(gnn_bgrad(WIDTH, N, dh.as<float>(), P, NPARTS, gB(r2(r)), s));
            HIPCK(gnn_gemm(WIDTH, WIDTH, N, dh.as<float>(), WIDTH, nullptr, WTb(r2(r)), nullptr, dT1.as<float>(), nullptr, s));
            HIPCK(gnn_relu_bwd(N * WIDTH, B[r].as<float>(), dT1.as<float>(), dT2.as<float>(), s));
            HIPCK(gnn_wgrad(2 * WIDTH, WIDTH, N, h[r].as<float>(), WIDTH, Mg[r].as<float>(), dT2.as<float>(), P, NPARTS,
                            gW(r1(r)), s));
            HIPCK(gnn_bgrad(WIDTH, N, dT2.as<float>(), P, NPARTS, gB(r1(r)), s));
            HIPCK(gnn_gemm(WIDTH, 2 * WIDTH, N, dT2.as<float>(), WIDTH, nullptr, WTb(r1(r)), nullptr, dU.as<float>(), nullptr, s));
            // residual path: dh_r = dh_{r+1} + dU[:, :32] + mean^T dU[:, 32:]
            HIPCK(gnn_split_concat(N, dU.as<float>(), dh.as<float>(), dm.as<float>(), s));
            HIPCK(gnn_mean_bwd(N, ap, ai, dm.as<float>(), dh.as<float>(), s));
        }
        // encoder (h0 is linear in H1)
        HIPCK(gnn_wgrad(WIDTH, WIDTH, N, H1.as<float>(), WIDTH, nullptr, dh.as<float>(), P, NPARTS, gW(e2()), s));
        HIPCK(gnn_bgrad(WIDTH, N, dh.as<float>(), P, NPARTS, gB(e2()), s));
        HIPCK(gnn_gemm(WIDTH, WIDTH, N, dh.as<float>(), WIDTH, nullptr, WTb(e2()), nullptr, dT1.as<float>(), nullptr, s));
        HIPCK(gnn_relu_bwd(N * WIDTH, A1.as<float>(), dT1.as<float>(), dT2.as<float>(), s));
        HIPCK(gnn_wgrad(FIN, WIDTH, N, feat.as<float>(), FIN, nullptr, dT2.as<float>(), P, NPARTS, gW(e1()), s));
        HIPCK(gnn_bgrad(WIDTH, N, dT2.as<float>(), P, NPARTS, gB(e1()), s));
    }
};

namespace sfem {

// node graph: nodes sharing an element, ascending, no self loops (the
// element neighbours of each node bucketed by node, then sorted per node)
void ensure_graph(sfem_ctx* c)
{
    if (c->d_adj_ptr.p) return;
    const int64_t nn = c->nn, ne = c->ne;
    std::vector<int64_t> off(nn + 1, 0);
    for (int64_t t = 0; t < 4 * ne; t++) off[c->tets_h[t] + 1] += 3;
    for (int64_t i = 0; i < nn; i++) off[i + 1] += off[i];
    std::vector<int> nb(off[nn]);
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t t = 0; t < ne; t++) {
        const int32_t* v = &c->tets_h[4 * t];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                if (a != b) nb[fill[v[a]]++] = v[b];
    }
    std::vector<int> ptr(nn + 1, 0), idx;
    idx.reserve(nb.size());
    for (int64_t i = 0; i < nn; i++) {
        auto first = nb.begin() + off[i], last = nb.begin() + off[i + 1];
        std::sort(first, last);
        for (auto it = first; it != last; ++it)
            if (*it != (int)i && (it == first || *it != *(it - 1))) idx.push_back(*it);
        ptr[i + 1] = (int)idx.size();
    }
    c->d_adj_ptr.upload(ptr);
    c->d_adj_idx.upload(idx);
    c->nadj = (int64_t)idx.size();
}

void gnn_install(sfem_ctx* c)
{
    sfem_gnn* g = c->gnn;
    ensure_graph(c);
    g->forward(c, c->stream, false);
    if (!c->d_minv_learned.p || c->d_minv_learned.n < 9 * (size_t)c->nn * sizeof(float))
        c->d_minv_learned.alloc(9 * (size_t)c->nn * sizeof(float));
    HIPCK(hipMemcpyAsync(c->d_minv_learned.p, g->minv.p, 9 * (size_t)c->nn * sizeof(float), hipMemcpyDeviceToDevice,
                         c->stream));
}

}  // namespace sfem

namespace {

// loss (and dMinv when grad) over pairs [first, first + S) with the given
// blocks; the caller has run forward() (learned) or filled g->minv (Jacobi)
double batch_loss(sfem_gnn* g, sfem_ctx* c, const std::vector<int64_t>& ids, const float* minv, bool grad)
{
    const int S = (int)ids.size();
    const int64_t N = c->nn, n3 = 3 * N;
    hipStream_t s = c->stream;
    g->batch_space(c, S);
    for (int k = 0; k < S; k++)
        HIPCK(hipMemcpyAsync(g->rb.as<float>() + k * n3, c->pairs.r[ids[k]]->p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    HIPCK(gnn_apply_blocks(N, S, minv, g->rb.as<float>(), g->zb.as<float>(), s));
    for (int k = 0; k < S; k++) hessvec(c, g->zb.as<float>() + k * n3, g->qb.as<float>() + k * n3, nullptr);
    HIPCK(gnn_pair_dots(n3, S, g->zb.as<float>(), g->rb.as<float>(), g->qb.as<float>(), g->dpart.as<double>(), NPARTS, s));
    std::vector<double> part((size_t)NPARTS * S * 2);
    HIPCK(hipMemcpyAsync(part.data(), g->dpart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    double loss = 0;
    std::vector<float> c12(2 * MAX_BATCH, 0.f);
    for (int k = 0; k < S; k++) {
        double a = 0, b = 0;
        for (int p = 0; p < NPARTS; p++) {   // fixed order
            a += part[((size_t)p * S + k) * 2];
            b += part[((size_t)p * S + k) * 2 + 1];
        }
        const double er = c->pairs.er[ids[k]];
        loss += 1.0 - a * a / (b * er);
        const double f = 2.0 * a / (b * er) / S;
        c12[k] = (float)(f * a / b);
        c12[MAX_BATCH + k] = (float)f;
    }
    loss /= S;
    if (grad) {
        HIPCK(hipMemcpyAsync(g->c12.p, c12.data(), c12.size() * sizeof(float), hipMemcpyHostToDevice, s));
        HIPCK(gnn_loss_grad_blocks(N, S, g->rb.as<float>(), g->qb.as<float>(), g->c12.as<float>(),
                                   g->c12.as<float>() + MAX_BATCH, g->dMinv.as<float>(), s));
    }
    return loss;
}

void check_pairs(sfem_ctx* c, int64_t first, int count)
{
    if (count <= 0 || count > MAX_BATCH || first < 0 || first + count > c->pairs.n)
        throw std::runtime_error("pair range out of bounds or batch above 16");
}

// hook of the recording PCG run: snapshot r and r.(e - x)
struct Recorder {
    int max_snaps;
    int taken = 0;
    const float* e;
};

// runs between CG launches (recorder_hook): its partials have a buffer of
// their own, since the CG's own partials carry over to the next launch
void record_pair(sfem_ctx* c, const float* e, const float* x, const float* r)
{
    const int64_t n3 = 3 * c->nn;
    auto buf = std::make_unique<DevBuf>();
    buf->alloc(n3 * sizeof(float));
    HIPCK(hipMemcpyAsync(buf->p, r, n3 * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    double part[256];
    if (!c->pairs.d_part.p) c->pairs.d_part.alloc(sizeof part);
    HIPCK(gnn_err_dot(n3, e, x, r, c->pairs.d_part.as<double>(), 256, c->stream));
    HIPCK(hipMemcpyAsync(part, c->pairs.d_part.p, sizeof part, hipMemcpyDeviceToHost, c->stream));
    HIPCK(hipStreamSynchronize(c->stream));
    double er = 0;
    for (double v : part) er += v;
    c->pairs.r.push_back(std::move(buf));
    c->pairs.er.push_back(er);
    c->pairs.n++;
}

void recorder_hook(sfem_ctx* c, int, void* user)
{
    Recorder* rec = (Recorder*)user;
    if (rec->taken >= rec->max_snaps) return;
    record_pair(c, rec->e, c->d_x32.as<float>(), c->d_r32.as<float>());
    rec->taken++;
}

}  // namespace

extern "C" {

int sfem_gnn_create(uint64_t seed, sfem_gnn** out)
{
    if (!out) return fail(SFEM_EINVAL, "sfem_gnn_create: null argument");
    *out = nullptr;
    sfem_gnn* g = new sfem_gnn;
    int rc = guard_call("sfem_gnn_create", [&]() -> int {
        g->init(seed);
        return SFEM_OK;
    });
    if (rc != SFEM_OK) {
        delete g;
        return rc;
    }
    *out = g;
    return SFEM_OK;
}

void sfem_gnn_destroy(sfem_gnn* g) { delete g; }

int sfem_gnn_num_params(sfem_gnn* g, int64_t* n)
{
    if (!g || !n) return fail(SFEM_EINVAL, "sfem_gnn_num_params: null argument");
    *n = (int64_t)g->nparam;
    return SFEM_OK;
}

int sfem_gnn_get_params(sfem_gnn* g, float* w)
{
    if (!g || !w) return fail(SFEM_EINVAL, "sfem_gnn_get_params: null argument");
    return guard_call("sfem_gnn_get_params", [&]() -> int {
        HIPCK(hipStreamSynchronize(g->stream));
        HIPCK(hipMemcpy(w, g->w.p, g->nparam * sizeof(float), hipMemcpyDeviceToHost));
        return SFEM_OK;
    });
}

int sfem_gnn_set_params(sfem_gnn* g, const float* w)
{
    if (!g || !w) return fail(SFEM_EINVAL, "sfem_gnn_set_params: null argument");
    return guard_call("sfem_gnn_set_params", [&]() -> int {
        HIPCK(hipMemcpy(g->w.p, w, g->nparam * sizeof(float), hipMemcpyHostToDevice));
        g->refresh_bf16();
        HIPCK(hipStreamSynchronize(g->stream));
        return SFEM_OK;
    });
}

int sfem_set_learned_preconditioner(sfem_ctx* c, sfem_gnn* g)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_set_learned_preconditioner: null context");
    c->gnn = g;
    if (!g) return SFEM_OK;
    // the node graph is part of the mesh set-up: built here, not in a solve
    return guard_call("sfem_set_learned_preconditioner", [&]() -> int {
        ensure_graph(c);
        return SFEM_OK;
    });
}

int sfem_collect_pairs(sfem_ctx* c, const double* load, int nsnap, int snap_every, double rtol, int* collected)
{
    if (!c || !load || nsnap <= 0 || snap_every <= 0 || !(rtol > 0))
        return fail(SFEM_EINVAL, "sfem_collect_pairs: bad argument");
    return guard_call("sfem_collect_pairs", [&]() -> int {
        const int64_t n3 = 3 * c->nn;
        std::vector<float> b(n3);
        for (int64_t i = 0; i < c->nn; i++)
            for (int k = 0; k < 3; k++) b[3 * i + k] = c->fixed_h[i] ? 0.f : (float)load[3 * i + k];
        build_jacobi(c);
        // 1. the solution e of K e = load (block-Jacobi PCG to rtol)
        HIPCK(hipMemcpyAsync(c->d_b32.p, b.data(), n3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        CgResult cg = pcg(c, c->d_minv.as<float>(), 1, 200000, 1, rtol * rtol, 1, 64, nullptr);
        if (cg.status != 1) throw std::runtime_error("pair collection: solve did not converge");
        HIPCK(hipMemcpyAsync(c->d_d32.p, c->d_x32.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        // 2. the same PCG again, snapshotting (r_k, r_k . (e - x_k)); K (e - x_k) = r_k
        const int before = (int)c->pairs.n;
        HIPCK(hipMemsetAsync(c->d_q32.p, 0, n3 * sizeof(float), c->stream));   // x0 = 0
        record_pair(c, c->d_d32.as<float>(), c->d_q32.as<float>(), c->d_b32.as<float>());
        Recorder rec{nsnap - 1, 0, c->d_d32.as<float>()};
        if (rec.max_snaps > 0)
            pcg(c, c->d_minv.as<float>(), 1, rec.max_snaps * snap_every, 1, rtol * rtol, 1, snap_every, nullptr,
                recorder_hook, &rec);
        if (collected) *collected = (int)(c->pairs.n - before);
        return SFEM_OK;
    });
}

int sfem_collect_rough_pairs(sfem_ctx* c, int count, uint64_t seed, int* collected)
{
    if (!c || count <= 0) return fail(SFEM_EINVAL, "sfem_collect_rough_pairs: bad argument");
    return guard_call("sfem_collect_rough_pairs", [&]() -> int {
        const int64_t n3 = 3 * c->nn;
        std::vector<float> e(n3);
        HIPCK(hipMemsetAsync(c->d_x32.p, 0, n3 * sizeof(float), c->stream));
        for (int k = 0; k < count; k++) {
            // e: independent uniform [-1, 1) per free dof (splitmix64 of seed, pair, dof), 0 at fixed nodes
            for (int64_t i = 0; i < n3; i++) {
                uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)k * 0xBF58476D1CE4E5B9ull + (uint64_t)i;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                e[i] = c->fixed_h[i / 3] ? 0.f : (float)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
            }
            HIPCK(hipMemcpyAsync(c->d_d32.p, e.data(), n3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
            hessvec(c, c->d_d32.as<float>(), c->d_q32.as<float>(), nullptr);
            record_pair(c, c->d_d32.as<float>(), c->d_x32.as<float>(), c->d_q32.as<float>());
        }
        if (collected) *collected = count;
        return SFEM_OK;
    });
}

int sfem_num_pairs(sfem_ctx* c, int64_t* n)
{
    if (!c || !n) return fail(SFEM_EINVAL, "sfem_num_pairs: null argument");
    *n = c->pairs.n;
    return SFEM_OK;
}

int sfem_get_pair(sfem_ctx* c, int64_t k, float* r, double* er)
{
    if (!c || k < 0 || k >= c->pairs.n) return fail(SFEM_EINVAL, "sfem_get_pair: bad index");
    return guard_call("sfem_get_pair", [&]() -> int {
        if (r) HIPCK(hipMemcpy(r, c->pairs.r[k]->p, 3 * c->nn * sizeof(float), hipMemcpyDeviceToHost));
        if (er) *er = c->pairs.er[k];
        return SFEM_OK;
    });
}

int sfem_clear_pairs(sfem_ctx* c)
{
    if (!c) return fail(SFEM_EINVAL, "sfem_clear_pairs: null context");
    c->pairs.r.clear();
    c->pairs.er.clear();
    c->pairs.n = 0;
    return SFEM_OK;
}

int sfem_gnn_loss(sfem_gnn* g, sfem_ctx* c, int64_t first, int count, double* loss, float* grad)
{
    if (!g || !c || !loss) return fail(SFEM_EINVAL, "sfem_gnn_loss: null argument");
    return guard_call("sfem_gnn_loss", [&]() -> int {
        check_pairs(c, first, count);
        ensure_graph(c);
        build_jacobi(c);
        g->forward(c, c->stream);
        std::vector<int64_t> ids(count);
        for (int k = 0; k < count; k++) ids[k] = first + k;
        *loss = batch_loss(g, c, ids, g->minv.as<float>(), grad != nullptr);
        if (grad) {
            g->backward(c, c->stream);
            HIPCK(hipMemcpyAsync(grad, g->g.p, g->nparam * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        }
        HIPCK(hipStreamSynchronize(c->stream));
        return SFEM_OK;
    });
}

int sfem_gnn_eval(sfem_gnn* g, sfem_ctx* c, double* loss_learned, double* loss_jacobi)
{
    if (!g || !c) return fail(SFEM_EINVAL, "sfem_gnn_eval: null argument");
    return guard_call("sfem_gnn_eval", [&]() -> int {
        if (c->pairs.n == 0) throw std::runtime_error("no pairs collected");
        ensure_graph(c);
        build_jacobi(c);
        g->forward(c, c->stream);
        double ll = 0, lj = 0;
        int64_t done = 0;
        while (done < c->pairs.n) {
            const int S = (int)std::min<int64_t>(MAX_BATCH, c->pairs.n - done);
            std::vector<int64_t> ids(S);
            for (int k = 0; k < S; k++) ids[k] = done + k;
            ll += S * batch_loss(g, c, ids, g->minv.as<float>(), false);
            lj += S * batch_loss(g, c, ids, c->d_minv.as<float>(), false);
            done += S;
        }
        if (loss_learned) *loss_learned = ll / c->pairs.n;
        if (loss_jacobi) *loss_jacobi = lj / c->pairs.n;
        return SFEM_OK;
    });
}

int sfem_gnn_train(sfem_gnn* g, sfem_ctx* c, int steps, int batch, float lr, uint64_t seed, double* losses)
{
    if (!g || !c || steps <= 0 || batch <= 0 || batch > MAX_BATCH || !(lr > 0))
        return fail(SFEM_EINVAL, "sfem_gnn_train: bad argument");
    return guard_call("sfem_gnn_train", [&]() -> int {
        if (c->pairs.n == 0) throw std::runtime_error("no pairs collected");
        ensure_graph(c);
        build_jacobi(c);
        uint64_t s = seed;
        const float b1 = 0.9f, b2 = 0.999f;
        for (int it = 0; it < steps; it++) {
            std::vector<int64_t> ids(batch);
            for (int k = 0; k < batch; k++) ids[k] = (int64_t)(splitmix(s) % (uint64_t)c->pairs.n);
            g->forward(c, c->stream);
            const double l = batch_loss(g, c, ids, g->minv.as<float>(), true);
            g->backward(c, c->stream);
            g->t++;
            const float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)g->t)) / (1.f - std::pow(b1, (float)g->t));
            HIPCK(gnn_adam((int)g->nparam, g->w.as<float>(), g->g.as<float>(), g->m.as<float>(), g->v.as<float>(), lr_t, b1,
                           b2, 1e-8f, 1e3f, c->stream));
            for (const Layer& l2 : g->L)
                HIPCK(gnn_to_bf16_pair(l2.K, l2.NO, g->W(l2), g->wb.as<uint16_t>() + l2.wb, g->wtb.as<uint16_t>() + l2.wb,
                                       c->stream));
            if (losses) losses[it] = l;
        }
        HIPCK(hipStreamSynchronize(c->stream));
        return SFEM_OK;
    });
}

int sfem_gnn_inputs(sfem_gnn* g, sfem_ctx* c, float* feat, float* cbase, float* minv)
{
    if (!g || !c) return fail(SFEM_EINVAL, "sfem_gnn_inputs: null argument");
    return guard_call("sfem_gnn_inputs", [&]() -> int {
        ensure_graph(c);
        build_jacobi(c);
        g->forward(c, c->stream);
        const size_t N = c->nn;
        if (feat) HIPCK(hipMemcpyAsync(feat, g->feat.p, N * FIN * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if (cbase) HIPCK(hipMemcpyAsync(cbase, g->cbase.p, N * 6 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if (minv) HIPCK(hipMemcpyAsync(minv, g->minv.p, N * 9 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCK(hipStreamSynchronize(c->stream));
        return SFEM_OK;
    });
}

int sfem_gnn_gemm_check(sfem_gnn* g, int layer, int64_t N, const float* X, float* Y)
{
    if (!g || !X || !Y || layer < 0 || layer >= (int)g->L.size()) return fail(SFEM_EINVAL, "sfem_gnn_gemm_check: bad argument");
    return guard_call("sfem_gnn_gemm_check", [&]() -> int {
        const Layer& l = g->L[layer];
        DevBuf x, y;
        x.alloc(N * l.K * sizeof(float));
        y.alloc(N * l.NO * sizeof(float));
        HIPCK(hipMemcpy(x.p, X, N * l.K * sizeof(float), hipMemcpyHostToDevice));
        HIPCK(gnn_gemm(l.K, l.NO, N, x.as<float>(), l.K, nullptr, g->Wb(l), g->Bias(l), y.as<float>(), nullptr, g->stream));
        HIPCK(hipStreamSynchronize(g->stream));
        HIPCK(hipMemcpy(Y, y.p, N * l.NO * sizeof(float), hipMemcpyDeviceToHost));
        return SFEM_OK;
    });
}

}  // extern "C"
