// api.hip -- the C ABI (include/ann_gpu.h): context, potential upload, batch,
// entry points, kernel timing.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "net_dev.h"

namespace ann {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

static int pool_event(ann_ctx* ctx) {
    if (ctx->evused == (int)ctx->evpool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->evpool.push_back(e);
    }
    return ctx->evused++;
}

static bool timed(const ann_ctx* ctx, const char* name) {
    return ctx->prof && !ctx->capturing && (ctx->prof_only.empty() || ctx->prof_only == name);
}
Launch::Launch(ann_ctx* c, const char* n) : ctx(c), name(n) {
    if (!timed(ctx, name)) return;
    const int a = pool_event(ctx), b = pool_event(ctx);
    if (a < 0 || b < 0) return;
    hipEventRecord(ctx->evpool[a], ctx->stream);
    ctx->timers[name].ev.push_back({a, b});
}
Launch::~Launch() {
    if (!timed(ctx, name)) return;
    auto& v = ctx->timers[name].ev;
    if (!v.empty()) hipEventRecord(ctx->evpool[v.back().second], ctx->stream);
}

// resolve every recorded launch into the per-name totals and recycle the events
static int drain_timers(ann_ctx* ctx) {
    ANN_HIP(hipStreamSynchronize(ctx->stream));
    for (auto& kv : ctx->timers) {
        Timer& T = kv.second;
        for (auto& e : T.ev) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ctx->evpool[e.first], ctx->evpool[e.second]) == hipSuccess) {
                T.total_ms += ms;
                T.count += 1;
            }
        }
        T.ev.clear();
    }
    ctx->evused = 0;
    return ANN_OK;
}

hipError_t struc_index(ann_ctx* ctx);
bool fast_shape_supported(int ng, int nk);
bool vec_net_supported(const ann_ctx* ctx);
bool mfma_net_supported(const DevModel& M);
hipError_t train_step_impl(ann_ctx* ctx, int form, const double* Eref, const double* Fref, double wE, double wF,
                           double gamma, double* loss, double* grad);

hipError_t staged_upload(ann_ctx* ctx, int slot, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    ann_ctx::Staged& s = ctx->staged[slot];
    hipError_t e;
    if (s.pending) {
        if ((e = hipEventSynchronize(s.ev)) != hipSuccess) return e;
        s.pending = false;
    }
    if (bytes > s.cap) {
        if (s.p) {
            e = hipHostFree(s.p);
            s.p = nullptr;
            s.cap = 0;
            if (e != hipSuccess) return e;
        }
        const size_t c = bytes + bytes / 4;
        if ((e = hipHostMalloc(&s.p, c)) != hipSuccess) return e;
        s.cap = c;
    }
    if (!s.ev && (e = hipEventCreateWithFlags(&s.ev, hipEventDisableTiming)) != hipSuccess) return e;
    std::memcpy(s.p, src, bytes);
    if ((e = hipMemcpyAsync(dst, s.p, bytes, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) return e;
    if ((e = hipEventRecord(s.ev, ctx->stream)) != hipSuccess) return e;
    s.pending = true;
    return hipSuccess;
}

template <typename T>
static hipError_t dalloc(T*& p, size_t n) {
    if (p) { hipFree(p); p = nullptr; }
    return hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1));
}

}  // namespace ann

using namespace ann;

#define CHECK_CTX(c) \
    if (!(c)) return fail(ANN_ERR_ARG, "null context")

extern "C" const char* ann_last_error(void) { return g_err.c_str(); }

// Inputs whose scaled value s = inpw (G + inpb) an FP32 G cannot resolve: with
// |inpw inpb| = R, the ~1e-7 relative rounding of G becomes ~1e-7 R in s.
// Those above ctx->pre_ratio are recomputed in FP64 and centred (precise.hip);
// the network applies inpb_net (inpb with those entries zeroed) to rows whose
// flagged entries hold G + inpb.  Uploads the FP32 and FP64 scaling, the items
// and the flags.
static int update_precise(ann_ctx* ctx) {
    DevModel& M = ctx->hm;
    const size_t tot = ctx->h_inpw.size();
    std::vector<PreItem> items;
    std::vector<uint8_t> flag(std::max<size_t>(tot, 1), 0);
    std::vector<float> iw(tot), ib(tot), ibn(tot);
    for (int a = 0; a < M.S; ++a) {
        M.pre_off[a] = (int)items.size();
        const NetDesc& D = M.net[a];
        for (int k = 0; k < D.nin; ++k) {
            const double w = ctx->h_inpw[D.inoff + k], b = ctx->h_inpb[D.inoff + k];
            iw[D.inoff + k] = (float)w;
            ib[D.inoff + k] = (float)b;
            const bool f = ctx->pre_ratio < 0 || std::fabs(w * b) > ctx->pre_ratio;
            ibn[D.inoff + k] = f ? 0.0f : (float)b;
            if (!f) continue;
            flag[D.inoff + k] = 1;
            PreItem it{k, -1, 0, 0, 0, 0, 0, 0};
            for (int c = 0; c < M.S && it.mode < 0; ++c) {
                const RadBlock& rb = M.rad[a * M.S + c];
                if (k >= rb.goff && k < rb.goff + rb.n) { it.mode = 0; it.blk = c; it.fn = k - rb.goff; }
            }
            for (int p = 0; p < M.NP && it.mode < 0; ++p) {
                const AngBlock& ab = M.ang[a * M.NP + p];
                if (k < ab.goff || k >= ab.goff + ab.n) continue;
                it.mode = 2;
                it.blk = p;
                it.fn = k - ab.goff;
                if (M.moment_z > 0) {       // every block is a GAUSS grid with integer zeta <= kMomZ
                    for (int g = 0; g < ab.ng; ++g)
                        for (int q = 0; q < ab.nk; ++q)
                            if (ab.outm[g * ab.nk + q] == it.fn) {
                                it.mode = 1;
                                it.u = ab.gu[g];
                                it.zint = ab.zint[q];
                                it.lam = ab.lamh[q] > 0 ? 1 : (ab.lamh[q] < 0 ? -1 : 0);
                            }
                }
            }
            if (it.mode < 0) return fail(ANN_ERR_ARG, "input outside every basis block");
            items.push_back(it);
        }
    }
    // when more than half of the inputs need FP64, all of them are evaluated in
    // FP64: the FP64 kernel's cost is per neighbour rather than per input, and
    // the FP32 basis pass is then skipped altogether
    if (ctx->pre_ratio >= 0 && 2 * items.size() > tot) {
        const double keep = ctx->pre_ratio;
        ctx->pre_ratio = -1.0;
        const int rc = update_precise(ctx);
        ctx->pre_ratio = keep;
        return rc;
    }
    M.pre_off[M.S] = (int)items.size();
    for (int a = M.S + 1; a <= kMaxSpecies; ++a) M.pre_off[a] = (int)items.size();
    ctx->npre = (int)items.size();
    ctx->nin_total = (int)tot;
    hipStream_t st = ctx->stream;
    if (tot > 0) {
        ANN_HIP(hipMemcpyAsync(ctx->da.inpw, iw.data(), sizeof(float) * tot, hipMemcpyHostToDevice, st));
        ANN_HIP(hipMemcpyAsync(ctx->da.inpb, ib.data(), sizeof(float) * tot, hipMemcpyHostToDevice, st));
        ANN_HIP(hipMemcpyAsync(ctx->da.inpb_net, ibn.data(), sizeof(float) * tot, hipMemcpyHostToDevice, st));
        ANN_HIP(hipMemcpyAsync(ctx->da.inpb64, ctx->h_inpb.data(), sizeof(double) * tot, hipMemcpyHostToDevice, st));
        ANN_HIP(hipMemcpyAsync(ctx->da.preflag, flag.data(), tot, hipMemcpyHostToDevice, st));
    }
    if (!items.empty())
        ANN_HIP(hipMemcpyAsync(ctx->da.pitems, items.data(), sizeof(PreItem) * items.size(), hipMemcpyHostToDevice, st));
    ANN_HIP(hipMemcpyAsync(ctx->dm, &M, sizeof(DevModel), hipMemcpyHostToDevice, st));
    ANN_HIP(hipStreamSynchronize(st));
    ctx->scaling_gen++;
    ctx->G_valid = false;          // cached centred rows depend on the scaling
    return ANN_OK;
}

static int upload_model(ann_ctx* ctx, const ann_model_desc* d) {
    const int S = d->nspecies;
    if (S < 1 || S > kMaxSpecies) return fail(ANN_ERR_UNSUPPORTED, "number of species must be in [1, 8]");
    if (!(d->rc > 0)) return fail(ANN_ERR_ARG, "invalid cutoff radius");
    if (d->cutoff < 1 || d->cutoff > 4) return fail(ANN_ERR_ARG, "invalid cutoff name");
    const int NP = S * (S + 1) / 2;
    DevModel& M = ctx->hm;
    std::memset(&M, 0, sizeof(M));
    M.S = S;
    M.NP = NP;
    M.cutoff = d->cutoff;
    M.rc = (float)d->rc;
    M.rci = (float)(1.0 / d->rc);
    M.pirc = (float)(3.141592653589793 / d->rc);
    std::vector<float> rrs, reta, aeta, azeta, alam;
    std::vector<double> rrs64, reta64, aeta64, azeta64;   // the same tables in FP64 (FP64-evaluated inputs)
    ctx->rc64 = d->rc;
    // radial blocks
    std::vector<int> nin(S, 0);
    for (int a = 0; a < S; ++a) {
        int goff = 0;
        for (int b = 0; b < S; ++b) {
            const int t = a * S + b, n = d->rad_n[t];
            if (n < 0 || n > kMaxFunc) return fail(ANN_ERR_UNSUPPORTED, "radial basis larger than 32 functions");
            if (n > 0 && (d->rad_name[t] < 1 || d->rad_name[t] > 6)) return fail(ANN_ERR_ARG, "invalid radial function");
            RadBlock& rb = M.rad[t];
            rb.name = d->rad_name[t];
            rb.n = n;
            rb.goff = goff;
            rb.par = (int)rrs.size();
            for (int m = 0; m < n; ++m) {
                rrs.push_back((float)d->rad_rs[d->rad_off[t] + m]);
                reta.push_back((float)d->rad_eta[d->rad_off[t] + m]);
                rrs64.push_back(d->rad_rs[d->rad_off[t] + m]);
                reta64.push_back(d->rad_eta[d->rad_off[t] + m]);
            }
            goff += n;
        }
        nin[a] = goff;
    }
    // angular blocks
    M.all_grid = 1;
    for (int a = 0; a < S; ++a) {
        std::vector<double> uni;
        for (int p = 0; p < NP; ++p) {
            const int t = a * NP + p, n = d->ang_n[t];
            if (n < 0 || n > kMaxFunc) return fail(ANN_ERR_UNSUPPORTED, "angular basis larger than 32 functions");
            if (n > 0 && (d->ang_name[t] < 1 || d->ang_name[t] > 6)) return fail(ANN_ERR_ARG, "invalid angular function");
            AngBlock& ab = M.ang[t];
            ab.name = d->ang_name[t];
            ab.n = n;
            ab.goff = nin[a];
            ab.par = (int)aeta.size();
            const double* eta = d->ang_eta + d->ang_off[t];
            const double* zeta = d->ang_zeta + d->ang_off[t];
            const int* lam = d->ang_lambda + d->ang_off[t];
            for (int m = 0; m < n; ++m) {
                if (!(eta[m] > 0)) return fail(ANN_ERR_ARG, "angular width must be positive");
                aeta.push_back((float)eta[m]);
                azeta.push_back((float)zeta[m]);
                aeta64.push_back(eta[m]);
                azeta64.push_back(zeta[m]);
                alam.push_back((float)lam[m]);
            }
            nin[a] += n;
            if (n == 0) { ab.grid = 1; continue; }
            // grid detection: functions = {width g} x {(lambda, zeta) k}
            std::vector<double> ge;
            std::vector<std::pair<int, double>> ks;
            std::vector<int> gi(n), ki(n);
            for (int m = 0; m < n; ++m) {
                auto it = std::find(ge.begin(), ge.end(), eta[m]);
                gi[m] = (int)(it - ge.begin());
                if (it == ge.end()) ge.push_back(eta[m]);
                const std::pair<int, double> key(lam[m], zeta[m]);
                auto kt = std::find(ks.begin(), ks.end(), key);
                ki[m] = (int)(kt - ks.begin());
                if (kt == ks.end()) ks.push_back(key);
            }
            const int ng = (int)ge.size(), nk = (int)ks.size();
            // every family runs the triplet tiles (GAUSS with per-neighbour factors,
            // the others with the pair factor of r_j^2 + r_k^2); the moment form needs GAUSS
            bool grid = ng <= kGridG && nk <= kGridK && ng * nk == n && fast_shape_supported(ng, nk);
            if (grid) {
                std::vector<int> seen(ng * nk, 0);
                for (int m = 0; m < n; ++m) {
                    if (seen[gi[m] * nk + ki[m]]++) grid = false;
                    ab.outm[gi[m] * nk + ki[m]] = m;
                }
                for (int k = 0; k < nk; ++k) {
                    ab.lamh[k] = (float)(0.5 * ks[k].first);
                    ab.zeta[k] = (float)ks[k].second;
                    const double z = ks[k].second;
                    ab.zint[k] = (z == std::floor(z) && z >= 1 && z <= 64) ? (int)z : -1;
                    if (!(z >= 1.0)) grid = false;
                }
                for (int g = 0; g < ng && grid; ++g) {
                    const double ie = 1.0 / (ge[g] * ge[g]);
                    auto it = std::find(uni.begin(), uni.end(), ie);
                    if (it == uni.end()) {
                        if ((int)uni.size() >= kMaxUnionEta) { grid = false; break; }
                        uni.push_back(ie);
                        it = uni.end() - 1;
                    }
                    ab.gu[g] = (int)(it - uni.begin());
                    // the fast kernels address width g of a block as union width g
                    if (ab.gu[g] != g) grid = false;
                }
            }
            ab.grid = grid ? 1 : 0;
            ab.ng = ng;
            ab.nk = nk;
            if (!grid) M.all_grid = 0;
        }
        M.nunion[a] = (int)uni.size();
        for (size_t u = 0; u < uni.size(); ++u) {
            M.uieta2[a][u] = (float)uni[u];
            M.uieta2_64[a][u] = uni[u];
        }
        M.nin[a] = nin[a];
    }
    // the fast kernels are compiled per grid shape: every non-empty block must share it
    M.fast_ng = M.fast_nk = 0;
    for (int t = 0; t < S * NP; ++t) {
        const AngBlock& ab = M.ang[t];
        if (ab.n == 0) continue;
        if (M.fast_ng == 0) { M.fast_ng = ab.ng; M.fast_nk = ab.nk; }
        else if (ab.ng != M.fast_ng || ab.nk != M.fast_nk) M.all_grid = 0;
    }
    if (M.fast_ng == 0) { M.fast_ng = 1; M.fast_nk = 1; }
    // moment form: grid blocks with integer zeta in [1, kMomZ] and |lambda| <= 1
    M.moment_z = 0;
    if (M.all_grid) {
        bool ok = true;
        int z = 0;
        for (int t = 0; t < S * NP; ++t) {
            const AngBlock& ab = M.ang[t];
            if (ab.n == 0) continue;
            if (ab.name != 1) ok = false;                        // only GAUSS factorises
            for (int k = 0; k < ab.nk; ++k) {
                if (ab.zint[k] < 1 || ab.zint[k] > kMomZ || std::fabs(ab.lamh[k]) > 0.5f) ok = false;
                z = std::max(z, ab.zint[k]);
            }
        }
        if (ok && z > 0) M.moment_z = z;
    }
    for (int t = 0; t < S * NP; ++t) {
        AngBlock& ab = M.ang[t];
        for (int k = 0; k < kGridK; ++k)
            for (int m = 0; m <= kMomZ; ++m) {
                double c = 0.0;
                if (M.moment_z > 0 && k < ab.nk && m <= ab.zint[k]) {
                    c = std::pow(2.0, m - ab.zint[k]) * std::pow((double)ab.lamh[k], m);
                    for (int q = 0; q < m; ++q) c = c * (ab.zint[k] - q) / (q + 1);
                }
                ab.mcoef[k][m] = (float)c;
            }
    }
    int maxin = 1;
    for (int a = 0; a < S; ++a) maxin = std::max(maxin, nin[a]);
    ctx->max_nin = maxin;
    ctx->ldg = (maxin + 3) / 4 * 4;
    M.ldg = ctx->ldg;
    // networks
    std::vector<float> W, iw, ib;
    int ptot = 0;
    for (int a = 0; a < S; ++a) {
        NetDesc& D = M.net[a];
        D.nin = nin[a];
        D.nlayer = d->nn_nlayer[a];
        if (D.nlayer < 1 || D.nlayer > kMaxLayers) return fail(ANN_ERR_UNSUPPORTED, "networks may have 1..4 layers");
        D.neuron = d->nn_neuron[a];
        if (!neuron_supported(D.neuron)) return fail(ANN_ERR_UNSUPPORTED, "unsupported neuron type");
        D.c = (float)d->nn_c[a];
        int prev = nin[a], cnt = 0;
        for (int l = 0; l < D.nlayer; ++l) {
            D.nodes[l] = d->nn_nodes[d->nn_nodes_off[a] + l];
            if (D.nodes[l] < 1 || D.nodes[l] > kMaxWidth) return fail(ANN_ERR_UNSUPPORTED, "layer width must be 1..128");
        }
        if (D.nodes[D.nlayer - 1] != 1) return fail(ANN_ERR_UNSUPPORTED, "network output size must be 1");
        D.poff = ptot;
        for (int l = 0; l < D.nlayer; ++l) { D.boff[l] = ptot + cnt; cnt += D.nodes[l]; }
        for (int l = 0; l < D.nlayer; ++l) { D.woff[l] = ptot + cnt; cnt += D.nodes[l] * prev; prev = D.nodes[l]; }
        D.nparam = cnt;
        for (int q = 0; q < cnt; ++q) {
            const double v = d->params[d->nn_par_off[a] + q];
            ctx->params.push_back(v);
            W.push_back((float)v);
        }
        ptot += cnt;
        D.inoff = (int)iw.size();
        for (int m = 0; m < nin[a]; ++m) {
            iw.push_back((float)d->inpw[d->nn_inp_off[a] + m]);
            ib.push_back((float)d->inpb[d->nn_inp_off[a] + m]);
            ctx->h_inpw.push_back(d->inpw[d->nn_inp_off[a] + m]);
            ctx->h_inpb.push_back(d->inpb[d->nn_inp_off[a] + m]);
        }
        D.outw = (float)d->outw[a];
        D.outb = (float)d->outb[a];
        D.evac = d->evac[a];
    }
    ctx->nparams = ptot;
    ctx->S = S;
    {
        // operand tables of the matrix-core network: for hidden layer l, W_l as
        // the right operand of the forward product (K = inputs, columns =
        // neurons) and W_l^T for the backward product (K = neurons, columns =
        // inputs); per (k block of 32, column tile of 16, BF16 term) one
        // 16-byte vector per lane
        int acc = 0;
        for (int a = 0; a < S; ++a) {
            NetDesc& D = M.net[a];
            for (int l = 0; l < D.nlayer - 1; ++l) {
                const int ni = l == 0 ? D.nin : D.nodes[l - 1], no = D.nodes[l];
                D.wsf[l] = acc;
                acc += (ni + 31) / 32 * ((no + 15) / 16) * 3 * kWave;
                D.wsb[l] = acc;
                acc += (no + 31) / 32 * ((ni + 15) / 16) * 3 * kWave;
            }
        }
        ctx->wsplit_n = acc;
        ctx->wsplit_valid = false;
        // tables of the vector-ALU network (layout in net_dev.h): the weight
        // rows of the hidden layers, their biases, the output neuron
        int cacc = 0, cmax = 0;
        for (int a = 0; a < S; ++a) {
            NetDesc& D = M.net[a];
            D.wc0 = cacc;
            for (int l = 0; l < kMaxLayers; ++l) D.wcf[l] = D.wcs[l] = D.wcb[l] = 0;
            for (int l = 0; l < D.nlayer - 1; ++l) {
                const int ni = l == 0 ? D.nin : D.nodes[l - 1], no = D.nodes[l];
                D.wcs[l] = vec_row_stride(no);
                D.wcf[l] = cacc;
                cacc += vec_pad4(ni * D.wcs[l]);
            }
            for (int l = 0; l < D.nlayer - 1; ++l) {
                D.wcb[l] = cacc;
                cacc += vec_pad4(D.nodes[l]);
            }
            const int nlast = D.nlayer > 1 ? D.nodes[D.nlayer - 2] : D.nin;
            D.wco = cacc;
            cacc += (nlast + 3) / 4 * 4 + 4;
            D.wcn = cacc - D.wc0;
            cmax = std::max(cmax, D.wcn);
        }
        // transposed rows W_l^T[j][k] of the hidden layers (odd stride)
        int tacc = 0;
        for (int a = 0; a < S; ++a) {
            NetDesc& D = M.net[a];
            for (int l = 0; l < kMaxLayers; ++l) D.wtf[l] = D.wts[l] = 0;
            for (int l = 0; l < D.nlayer - 1; ++l) {
                const int ni = l == 0 ? D.nin : D.nodes[l - 1], no = D.nodes[l];
                D.wts[l] = vec_row_stride(ni);
                D.wtf[l] = tacc;
                tacc += vec_pad4(no * D.wts[l]);
            }
        }
        ctx->wtrans_n = tacc;
        ctx->wchunk_n = cacc;
        ctx->wchunk_max = cmax;
        ctx->wchunk_valid = false;
    }
    // device copies
    ANN_HIP(hipMalloc(&ctx->dm, sizeof(DevModel)));
    ANN_HIP(hipMemcpy(ctx->dm, &M, sizeof(DevModel), hipMemcpyHostToDevice));
    auto up = [](float*& dst, const std::vector<float>& v) -> hipError_t {
        hipError_t e = hipMalloc(&dst, sizeof(float) * std::max<size_t>(v.size(), 1));
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(dst, v.data(), sizeof(float) * v.size(), hipMemcpyHostToDevice);
    };
    ANN_HIP(up(ctx->da.rad_rs, rrs));
    ANN_HIP(up(ctx->da.rad_eta, reta));
    ANN_HIP(up(ctx->da.ang_eta, aeta));
    ANN_HIP(up(ctx->da.ang_zeta, azeta));
    ANN_HIP(up(ctx->da.ang_lam, alam));
    ANN_HIP(up(ctx->da.W, W));
    ANN_HIP(hipMalloc(&ctx->d_wsplit, sizeof(uint4) * std::max(ctx->wsplit_n, 1)));
    ANN_HIP(hipMalloc(&ctx->d_wchunk, sizeof(float) * std::max(ctx->wchunk_n, 4)));
    ANN_HIP(hipMalloc(&ctx->d_wtrans, sizeof(float) * std::max(ctx->wtrans_n, 4)));
    ANN_HIP(up(ctx->da.inpw, iw));
    ANN_HIP(up(ctx->da.inpb, ib));
    ANN_HIP(up(ctx->da.inpb_net, ib));
    {
        auto up64 = [](double*& dst, const std::vector<double>& v) -> hipError_t {
            hipError_t e = hipMalloc(&dst, sizeof(double) * std::max<size_t>(v.size(), 1));
            if (e != hipSuccess || v.empty()) return e;
            return hipMemcpy(dst, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice);
        };
        ANN_HIP(up64(ctx->da.rad_rs64, rrs64));
        ANN_HIP(up64(ctx->da.rad_eta64, reta64));
        ANN_HIP(up64(ctx->da.ang_eta64, aeta64));
        ANN_HIP(up64(ctx->da.ang_zeta64, azeta64));
        ANN_HIP(up64(ctx->da.inpb64, ctx->h_inpb));
        const size_t nin_tot = std::max<size_t>(ctx->h_inpw.size(), 1);
        ANN_HIP(hipMalloc(&ctx->da.pitems, sizeof(PreItem) * nin_tot));
        ANN_HIP(hipMalloc(&ctx->da.preflag, nin_tot));
    }
    ANN_HIP(hipMalloc(&ctx->da.params, sizeof(double) * ptot));
    ANN_HIP(hipMemcpy(ctx->da.params, ctx->params.data(), sizeof(double) * ptot, hipMemcpyHostToDevice));
    ANN_HIP(hipMalloc(&ctx->da.adam_m, sizeof(double) * ptot));
    ANN_HIP(hipMalloc(&ctx->da.adam_v, sizeof(double) * ptot));
    ANN_HIP(hipMemset(ctx->da.adam_m, 0, sizeof(double) * ptot));
    ANN_HIP(hipMemset(ctx->da.adam_v, 0, sizeof(double) * ptot));
    ANN_HIP(hipMalloc(&ctx->d_flag, sizeof(int) * 4));
    ANN_HIP(hipMemset(ctx->d_flag, 0, sizeof(int) * 4));
    ANN_HIP(hipHostMalloc((void**)&ctx->h_rb, sizeof(int) * kReadbackWords, hipHostMallocCoherent));
    ANN_HIP(hipHostGetDevicePointer((void**)&ctx->d_rb, ctx->h_rb, 0));
    for (int q = 0; q < kReadbackWords; ++q) ctx->h_rb[q] = 0;
    ANN_HIP(hipMalloc(&ctx->d_adam_t, sizeof(int)));
    ANN_HIP(hipMemset(ctx->d_adam_t, 0, sizeof(int)));
    ANN_HIP(split_weights(ctx, nullptr));
    ANN_HIP(hipStreamSynchronize(ctx->stream));
    return update_precise(ctx);
}

extern "C" int ann_create(const ann_model_desc* desc, ann_ctx** out) {
    if (!desc || !out) return fail(ANN_ERR_ARG, "null argument");
    *out = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(ANN_ERR_HIP, "no HIP device");
    ann_ctx* ctx = new ann_ctx;
    ctx->device = dev;
    // the allocation-free entry point's scratch address on this device
    const hipError_t es = noalloc_init(dev);
    if (es != hipSuccess) {
        delete ctx;
        return fail(ANN_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(es));
    }
    const int rc = upload_model(ctx, desc);
    if (rc != ANN_OK) {
        ann_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return ANN_OK;
}

extern "C" int ann_destroy(ann_ctx* ctx) {
    if (!ctx) return ANN_OK;
    hipDeviceSynchronize();
    pair_free(ctx);
    void* ptrs[] = {ctx->d_off_own, ctx->d_cells_own, ctx->d_species_own, ctx->dm, ctx->da.rad_rs, ctx->da.rad_eta, ctx->da.ang_eta, ctx->da.ang_zeta, ctx->da.ang_lam,
                    ctx->da.W, ctx->da.inpw, ctx->da.inpb, ctx->da.params, ctx->da.adam_m, ctx->da.adam_v,
                    ctx->d_struc_of, ctx->d_cellinfo, ctx->d_cellmat, ctx->d_bin_of, ctx->d_shift, ctx->d_bin_cnt,
                    ctx->d_bin_start, ctx->d_bin_atoms, ctx->d_keys, ctx->d_cnt, ctx->d_flag, ctx->nl.off,
                    ctx->nl.idx, ctx->nl.img, ctx->nl.rev, ctx->nl.geo, ctx->nl.nsp, ctx->d_G, ctx->d_dEdG,
                    ctx->d_Eatom, ctx->d_P, ctx->d_u, ctx->d_v, ctx->d_F, ctx->d_E, ctx->d_gE, ctx->d_atomvir,
                    ctx->d_scr, ctx->d_gpart, ctx->d_rank, ctx->d_spcount, ctx->d_rkflag, ctx->d_rkpre, ctx->d_perm, ctx->d_tiles, ctx->d_spend, ctx->d_grad_tiles, ctx->d_gsplit,
                    ctx->d_chunks, ctx->d_chunk_off, ctx->d_part, ctx->d_pos_ref, ctx->d_keatom, ctx->d_md_flag,
                    ctx->d_bin_struc, ctx->d_binrec, ctx->d_adam_t, ctx->d_wsplit, ctx->d_wchunk, ctx->d_wtrans, ctx->d_T,
                    ctx->da.pitems, ctx->da.preflag, ctx->da.inpb64, ctx->da.inpb_net, ctx->da.rad_rs64,
                    ctx->da.rad_eta64, ctx->da.ang_eta64, ctx->da.ang_zeta64, ctx->d_items, ctx->d_qoff,
                    ctx->d_queue};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (ann_ctx::Staged& st : ctx->staged) {
        if (st.p) hipHostFree(st.p);
        if (st.ev) hipEventDestroy(st.ev);
    }
    if (ctx->h_rb) hipHostFree(ctx->h_rb);
    for (hipEvent_t e : ctx->evpool) hipEventDestroy(e);
    if (ctx->gexec) hipGraphExecDestroy(ctx->gexec);
    if (ctx->ef_exec) hipGraphExecDestroy(ctx->ef_exec);
    for (hipEvent_t e : ctx->gev)
        if (e) hipEventDestroy(e);
    if (ctx->gstream) hipStreamDestroy(ctx->gstream);
    delete ctx;
    return ANN_OK;
}

extern "C" int ann_set_stream(ann_ctx* ctx, void* s) {
    CHECK_CTX(ctx);
    ctx->stream = (hipStream_t)s;
    return ANN_OK;
}

extern "C" int ann_info(ann_ctx* ctx, int* S, int* max_nin, int* ldg, int* np) {
    CHECK_CTX(ctx);
    if (S) *S = ctx->S;
    if (max_nin) *max_nin = ctx->max_nin;
    if (ldg) *ldg = ctx->ldg;
    if (np) *np = ctx->nparams;
    return ANN_OK;
}

extern "C" int ann_set_params(ann_ctx* ctx, const double* p, int n) {
    CHECK_CTX(ctx);
    if (!p || n != ctx->nparams) return fail(ANN_ERR_ARG, "parameter count mismatch");
    ctx->params.assign(p, p + n);
    std::vector<float> W(p, p + n);
    ANN_HIP(hipMemcpyAsync(ctx->da.params, p, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    ANN_HIP(hipMemcpyAsync(ctx->da.W, W.data(), sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    ANN_HIP(split_weights(ctx, nullptr));
    ANN_HIP(hipStreamSynchronize(ctx->stream));
    ANN_HIP(hipMemsetAsync(ctx->d_adam_t, 0, sizeof(int), ctx->stream));
    ANN_HIP(hipMemsetAsync(ctx->da.adam_m, 0, sizeof(double) * n, ctx->stream));
    ANN_HIP(hipMemsetAsync(ctx->da.adam_v, 0, sizeof(double) * n, ctx->stream));
    return ANN_OK;
}

extern "C" int ann_get_params(ann_ctx* ctx, double* p, int n) {
    CHECK_CTX(ctx);
    if (!p || n != ctx->nparams) return fail(ANN_ERR_ARG, "parameter count mismatch");
    ANN_HIP(hipMemcpyAsync(p, ctx->da.params, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    ANN_HIP(hipStreamSynchronize(ctx->stream));
    return ANN_OK;
}

extern "C" int ann_get_optimizer_state(ann_ctx* ctx, double* m, double* v, int* step, int n) {
    CHECK_CTX(ctx);
    if (!m || !v || !step || n != ctx->nparams) return fail(ANN_ERR_ARG, "optimizer state: count mismatch");
    ANN_HIP(hipMemcpyAsync(m, ctx->da.adam_m, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    ANN_HIP(hipMemcpyAsync(v, ctx->da.adam_v, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    ANN_HIP(hipMemcpyAsync(step, ctx->d_adam_t, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ANN_HIP(hipStreamSynchronize(ctx->stream));
    return ANN_OK;
}

extern "C" int ann_set_optimizer_state(ann_ctx* ctx, const double* m, const double* v, int step, int n) {
    CHECK_CTX(ctx);
    if (!m || !v || n != ctx->nparams || step < 0) return fail(ANN_ERR_ARG, "optimizer state: bad arguments");
    ANN_HIP(hipMemcpyAsync(ctx->da.adam_m, m, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    ANN_HIP(hipMemcpyAsync(ctx->da.adam_v, v, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    ANN_HIP(hipMemcpyAsync(ctx->d_adam_t, &step, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ANN_HIP(hipStreamSynchronize(ctx->stream));
    return ANN_OK;
}

extern "C" int ann_set_scaling(ann_ctx* ctx, int a, const double* inpw, const double* inpb, double outw,
                               double outb) {
    CHECK_CTX(ctx);
    if (a < 0 || a >= ctx->S || !inpw || !inpb) return fail(ANN_ERR_ARG, "invalid species or null scaling");
    NetDesc& D = ctx->hm.net[a];
    for (int k = 0; k < D.nin; ++k) {
        if (!std::isfinite(inpw[k]) || !std::isfinite(inpb[k])) return fail(ANN_ERR_ARG, "non-finite input scaling");
        ctx->h_inpw[D.inoff + k] = inpw[k];
        ctx->h_inpb[D.inoff + k] = inpb[k];
    }
    D.outw = (float)outw;
    D.outb = (float)outb;
    return update_precise(ctx);
}

extern "C" int ann_set_precise_ratio(ann_ctx* ctx, double ratio) {
    CHECK_CTX(ctx);
    if (std::isnan(ratio)) return fail(ANN_ERR_ARG, "invalid ratio");
    ctx->pre_ratio = ratio;
    return update_precise(ctx);
}

extern "C" int ann_precise_count(ann_ctx* ctx, int* count) {
    CHECK_CTX(ctx);
    if (!count) return fail(ANN_ERR_ARG, "null argument");
    *count = ctx->npre;
    return ANN_OK;
}

extern "C" int ann_set_batch(ann_ctx* ctx, const ann_batch* b) {
    CHECK_CTX(ctx);
    if (!b || b->nstruc < 1 || !b->atom_off || !b->d_atom_off || !b->d_pos || !b->d_cells || !b->d_species)
        return fail(ANN_ERR_ARG, "invalid batch");
    const int S = b->nstruc;
    if (b->atom_off[0] != 0) return fail(ANN_ERR_ARG, "atom_off[0] must be 0");
    for (int s = 0; s < S; ++s)
        if (b->atom_off[s + 1] <= b->atom_off[s]) return fail(ANN_ERR_ARG, "every structure needs at least one atom");
    const int N = b->atom_off[S];
    ctx->nstruc = S;
    ctx->natoms = N;
    ctx->atom_off.assign(b->atom_off, b->atom_off + S + 1);
    ctx->d_atom_off = b->d_atom_off;
    ctx->d_pos = b->d_pos;
    ctx->d_cells = b->d_cells;
    ctx->d_species = b->d_species;
    if (N > ctx->atom_cap) {
        ANN_HIP(dalloc(ctx->d_struc_of, N));
        ANN_HIP(dalloc(ctx->d_bin_of, N));
        ANN_HIP(dalloc(ctx->d_shift, 3 * (size_t)N));
        ANN_HIP(dalloc(ctx->d_bin_atoms, N));
        ANN_HIP(dalloc(ctx->d_cnt, N + 1));
        ANN_HIP(dalloc(ctx->nl.off, N + 1));
        ctx->atom_cap = N;
    }
    // chunks of the per-structure sums
    std::vector<int> chunks, chunk_off(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        chunk_off[s] = (int)chunks.size() / 3;
        for (int a = b->atom_off[s]; a < b->atom_off[s + 1]; a += kSumChunk) {
            chunks.push_back(s);
            chunks.push_back(a);
            chunks.push_back(std::min(a + kSumChunk, b->atom_off[s + 1]));
        }
    }
    const int nch = (int)chunks.size() / 3;
    chunk_off[S] = nch;
    if (nch > ctx->chunk_cap || S + 1 > ctx->chunk_cap) {
        const int cap = std::max(nch, S + 1);
        ANN_HIP(dalloc(ctx->d_chunks, 3 * (size_t)cap));
        ANN_HIP(dalloc(ctx->d_chunk_off, (size_t)cap + 1));
        ANN_HIP(dalloc(ctx->d_part, (size_t)kSumQ * cap));
        ctx->chunk_cap = cap;
    }
    ctx->nchunks = nch;
    ANN_HIP(staged_upload(ctx, kStChunks, ctx->d_chunks, chunks.data(), sizeof(int) * chunks.size()));
    ANN_HIP(staged_upload(ctx, kStChunkOff, ctx->d_chunk_off, chunk_off.data(), sizeof(int) * (S + 1)));
    ANN_HIP(hipMemsetAsync(ctx->d_cnt, 0, sizeof(int) * (N + 1), ctx->stream));
    ANN_HIP(struc_index(ctx));
    ctx->items_valid = false;      // persistent-dispatch items: built on first use
    ctx->h_cells_valid = false;    // host copies: only with ann_set_batch_host
    ctx->h_counts_valid = false;
    ctx->have_batch = true;
    ctx->batch_gen++;
    ctx->have_nlist = false;
    ctx->grid_valid = false;
    ctx->gen++;
    ctx->G_valid = false;
    ctx->rank_valid = false;
    ctx->perm_valid = false;
    return ANN_OK;
}

extern "C" int ann_set_batch_host(ann_ctx* ctx, int nstruc, const int* atom_off, const double* d_pos,
                                  const double* cells, const int* species) {
    CHECK_CTX(ctx);
    if (nstruc < 1 || !atom_off || !d_pos || !cells || !species) return fail(ANN_ERR_ARG, "invalid batch");
    if (atom_off[0] != 0) return fail(ANN_ERR_ARG, "atom_off[0] must be 0");
    for (int s = 0; s < nstruc; ++s)
        if (atom_off[s + 1] <= atom_off[s]) return fail(ANN_ERR_ARG, "every structure needs at least one atom");
    const int N = atom_off[nstruc];
    std::vector<int> count(ctx->S, 0);
    for (int i = 0; i < N; ++i) {
        if (species[i] < 0 || species[i] >= ctx->S) return fail(ANN_ERR_ARG, "species index out of range");
        ++count[species[i]];
    }
    // the context's own device copies (grow-only; a free waits for their
    // previous readers), filled through the pinned staging without a wait
    if (nstruc + 1 > ctx->own_struc_cap) {
        ANN_HIP(dalloc(ctx->d_off_own, 2 * ((size_t)nstruc + 1)));
        ANN_HIP(dalloc(ctx->d_cells_own, 18 * (size_t)nstruc));
        ctx->own_struc_cap = 2 * (nstruc + 1);
    }
    if (N > ctx->own_atom_cap) {
        ANN_HIP(dalloc(ctx->d_species_own, 2 * (size_t)N));
        ctx->own_atom_cap = 2 * N;
    }
    ANN_HIP(staged_upload(ctx, kStAtomOff, ctx->d_off_own, atom_off, sizeof(int) * ((size_t)nstruc + 1)));
    ANN_HIP(staged_upload(ctx, kStCells, ctx->d_cells_own, cells, sizeof(double) * 9 * (size_t)nstruc));
    ANN_HIP(staged_upload(ctx, kStSpecies, ctx->d_species_own, species, sizeof(int) * (size_t)N));
    const ann_batch b{nstruc, atom_off, ctx->d_off_own, d_pos, ctx->d_cells_own, ctx->d_species_own};
    const int rc = ann_set_batch(ctx, &b);
    if (rc) return rc;
    ctx->h_cells.assign(cells, cells + 9 * (size_t)nstruc);
    ctx->h_species.assign(species, species + N);
    ctx->h_species_count = count;
    ctx->h_cells_valid = true;
    ctx->h_counts_valid = true;
    return ANN_OK;
}

extern "C" int ann_neighbors(ann_ctx* ctx, double rc) {
    CHECK_CTX(ctx);
    if (!ctx->have_batch) return fail(ANN_ERR_STATE, "no batch set");
    const double r = rc > 0 ? rc : ctx->rc64;
    ctx->have_nlist = false;      // the build rewrites the list arrays in place
    if (!(rc >= 0) || std::isinf(rc)) return fail(ANN_ERR_ARG, "invalid neighbour cutoff");
    ctx->nb_status = ANN_OK;
    hipError_t e = neighbors_build(ctx, rc > 0 ? rc : ctx->rc64);
    if (ctx->nb_status != ANN_OK) return ctx->nb_status;     // message already set
    if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("neighbour build: ") + hipGetErrorString(e));
    ctx->have_nlist = true;
    ctx->G_valid = false;
    ctx->nl_rc = r;
    ctx->pos_ref_valid = false;   // the skin reference is taken by the next MD step
    ctx->gen++;
    return ANN_OK;
}

extern "C" int ann_set_neighbors(ann_ctx* ctx, const int* off, const int* idx, const int* img, double rc) {
    CHECK_CTX(ctx);
    if (!ctx->have_batch) return fail(ANN_ERR_STATE, "no batch set");
    if (!off || (off[ctx->natoms] > 0 && (!idx || !img))) return fail(ANN_ERR_ARG, "null neighbour list");
    if (!(rc > 0) || std::isinf(rc)) return fail(ANN_ERR_ARG, "invalid neighbour cutoff");
    ctx->have_nlist = false;
    const int r = set_neighbors_host(ctx, off, idx, img, rc);
    if (r != ANN_OK) return r;
    ctx->have_nlist = true;
    ctx->G_valid = false;
    ctx->nl_rc = rc;
    ctx->pos_ref_valid = false;
    ctx->gen++;
    return ANN_OK;
}

extern "C" int ann_refresh_geometry(ann_ctx* ctx) {
    CHECK_CTX(ctx);
    if (!ctx->have_nlist) return fail(ANN_ERR_STATE, "build the neighbour list first (ann_neighbors)");
    hipError_t e = refresh_geometry(ctx);
    if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("refresh_geometry: ") + hipGetErrorString(e));
    ctx->G_valid = false;
    ctx->gen++;
    return ANN_OK;
}

extern "C" int ann_wrap_positions(ann_ctx* ctx) {
    CHECK_CTX(ctx);
    if (!ctx->have_batch) return fail(ANN_ERR_STATE, "no batch set");
    hipError_t e = wrap_positions(ctx, const_cast<double*>(ctx->d_pos));
    if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("wrap_positions: ") + hipGetErrorString(e));
    ctx->have_nlist = false;      // image vectors of the old list refer to the unwrapped positions
    ctx->G_valid = false;
    ctx->gen++;
    return ANN_OK;
}

extern "C" int ann_nlist_size(ann_ctx* ctx, int64_t* np, int* maxn) {
    CHECK_CTX(ctx);
    if (!ctx->have_nlist) return fail(ANN_ERR_STATE, "no neighbour list");
    int rev_err = 0;
    ANN_HIP(hipMemcpyAsync(&rev_err, ctx->d_flag + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ANN_HIP(hipStreamSynchronize(ctx->stream));
    if (rev_err) {
        ANN_HIP(hipMemsetAsync(ctx->d_flag + 2, 0, sizeof(int), ctx->stream));
        return fail(ANN_ERR_HIP, "neighbour list is not symmetric (reverse slot missing)");
    }
    if (np) *np = ctx->nl.npairs;
    if (maxn) *maxn = ctx->nl.maxn;
    return ANN_OK;
}

namespace ann {
__global__ void k_unpack_img(int64_t P, const uint32_t* img, int* out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P) return;
    int a, b, c;
    unpack_img(img[s], a, b, c);
    out[3 * s] = a;
    out[3 * s + 1] = b;
    out[3 * s + 2] = c;
}
}  // namespace ann

extern "C" int ann_nlist_copy(ann_ctx* ctx, int* off, int* idx, int* img, int* rev, float* geo) {
    CHECK_CTX(ctx);
    if (!ctx->have_nlist) return fail(ANN_ERR_STATE, "no neighbour list");
    const int64_t P = ctx->nl.npairs;
    hipStream_t st = ctx->stream;
    if (off) ANN_HIP(hipMemcpyAsync(off, ctx->nl.off, sizeof(int) * (ctx->natoms + 1), hipMemcpyDeviceToDevice, st));
    if (P > 0) {
        if (idx) ANN_HIP(hipMemcpyAsync(idx, ctx->nl.idx, sizeof(int) * P, hipMemcpyDeviceToDevice, st));
        if (rev) ANN_HIP(hipMemcpyAsync(rev, ctx->nl.rev, sizeof(int) * P, hipMemcpyDeviceToDevice, st));
        if (geo) ANN_HIP(hipMemcpyAsync(geo, ctx->nl.geo, sizeof(float4) * P, hipMemcpyDeviceToDevice, st));
        if (img) {
            k_unpack_img<<<(unsigned)((P + 255) / 256), 256, 0, st>>>(P, ctx->nl.img, img);
            ANN_HIP(hipGetLastError());
        }
    }
    return ANN_OK;
}

static int need_nlist(ann_ctx* ctx) {
    if (!ctx) return fail(ANN_ERR_ARG, "null context");
    if (!ctx->have_nlist) return fail(ANN_ERR_STATE, "build the neighbour list first (ann_neighbors)");
    return ANN_OK;
}

static int check_form(int form) {
    if (form != ANN_FORM_PLAIN && form != ANN_FORM_FAST && form != ANN_FORM_TRIPLET)
        return fail(ANN_ERR_ARG, "invalid form");
    return ANN_OK;
}

#define RET_HIP(expr, what)                                                                  \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(ANN_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int ann_symm_fwd(ann_ctx* ctx, int form, float* G) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!G) return fail(ANN_ERR_ARG, "null output");
    RET_HIP(symm_fwd(ctx, form, G), "symm_fwd");
    return ANN_OK;
}

extern "C" int ann_net_fwd(ann_ctx* ctx, int form, const float* G, float* Eatom, float* dEdG) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!G || !Eatom) return fail(ANN_ERR_ARG, "null argument");
    RET_HIP(net_fwd(ctx, form, G, Eatom, dEdG), "net_fwd");
    return ANN_OK;
}

extern "C" int ann_symm_bwd(ann_ctx* ctx, int form, const float* dEdG, float* P) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!dEdG || !P) return fail(ANN_ERR_ARG, "null argument");
    RET_HIP(symm_bwd(ctx, form, dEdG, P), "symm_bwd");
    return ANN_OK;
}

static hipError_t ensure_work(ann_ctx* ctx);

extern "C" int ann_symm_fwd64(ann_ctx* ctx, int form, double* G64) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!G64) return fail(ANN_ERR_ARG, "null output");
    RET_HIP(ensure_work(ctx), "allocation");
    ctx->G_valid = false;
    RET_HIP(symm_fwd(ctx, form, ctx->d_G), "symm_fwd");
    RET_HIP(widen_rows(ctx, ctx->d_G, G64), "symm_fwd64");
    return ANN_OK;
}

extern "C" int ann_net_fwd64(ann_ctx* ctx, int form, const double* G64, float* Eatom, float* dEdG) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!G64 || !Eatom) return fail(ANN_ERR_ARG, "null argument");
    RET_HIP(ensure_work(ctx), "allocation");
    ctx->G_valid = false;
    RET_HIP(centre_rows(ctx, G64, ctx->d_G), "net_fwd64");
    RET_HIP(net_fwd(ctx, form, ctx->d_G, Eatom, dEdG, true), "net_fwd");
    return ANN_OK;
}

static hipError_t ensure_work(ann_ctx* ctx) {
    const int64_t N = ctx->natoms, P = ctx->nl.npairs;
    if (N > ctx->work_atoms) {
        hipError_t e;
        if ((e = dalloc(ctx->d_G, (size_t)N * ctx->ldg)) != hipSuccess) return e;
        if ((e = dalloc(ctx->d_dEdG, (size_t)N * ctx->ldg)) != hipSuccess) return e;
        if ((e = dalloc(ctx->d_v, (size_t)N * ctx->ldg)) != hipSuccess) return e;
        if ((e = dalloc(ctx->d_Eatom, N)) != hipSuccess) return e;
        if ((e = dalloc(ctx->d_u, 3 * (size_t)N)) != hipSuccess) return e;
        if ((e = dalloc(ctx->d_F, 3 * (size_t)N)) != hipSuccess) return e;
        if ((e = dalloc(ctx->d_atomvir, 9 * (size_t)N)) != hipSuccess) return e;
        if (ctx->hm.moment_z > 0) {   // moments of the moment form: [N][S][4][monomials]
            const int z = ctx->hm.moment_z, nm = (z + 1) * (z + 2) * (z + 3) / 6;
            if ((e = dalloc(ctx->d_T, (size_t)N * ctx->S * kMaxUnionEta * nm)) != hipSuccess) return e;
        }
        ctx->T_gen = ~0ull;
        ctx->work_atoms = N;
    }
    if (P > ctx->work_pairs || !ctx->d_P) {
        hipError_t e;
        if ((e = dalloc(ctx->d_P, 3 * (size_t)std::max<int64_t>(P, 1))) != hipSuccess) return e;
        ctx->work_pairs = std::max<int64_t>(P, 1);
    }
    if (!ctx->d_E || ctx->nstruc > ctx->e_cap) {
        hipError_t e;
        if ((e = dalloc(ctx->d_E, ctx->nstruc)) != hipSuccess) return e;
        if ((e = dalloc(ctx->d_gE, ctx->nstruc)) != hipSuccess) return e;
        ctx->e_cap = ctx->nstruc;
    }
    return hipSuccess;
}

namespace ann {
hipError_t ensure_work_buffers(ann_ctx* ctx) { return ensure_work(ctx); }
}

extern "C" int ann_force_reduce(ann_ctx* ctx, const float* P, const float* Eatom, int with_evac, double* F,
                                double* E, double* V) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if (!P || (E && !Eatom)) return fail(ANN_ERR_ARG, "null argument");
    RET_HIP(ensure_work(ctx), "allocation");
    RET_HIP(force_reduce(ctx, P, Eatom, with_evac, F, E, V), "force_reduce");
    return ANN_OK;
}

extern "C" int ann_force_reduce_noalloc(ann_ctx* ctx, const float* P, const float* Eatom, int with_evac,
                                        double* F, double* E, double* V) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if (!P || (E && !Eatom)) return fail(ANN_ERR_ARG, "null argument");
    const int64_t need = noalloc_need(ctx, V != nullptr);
    if (need > kNoAllocScratch)
        return fail(ANN_ERR_CAPACITY, "ann_force_reduce_noalloc: the batch needs " + std::to_string(need) +
                                          " doubles of scratch; the static scratch holds " +
                                          std::to_string(kNoAllocScratch) + " (32 MiB; 9 per atom with the virial)");
    RET_HIP(force_reduce_noalloc(ctx, P, Eatom, with_evac, F, E, V), "force_reduce_noalloc");
    return ANN_OK;
}

static hipError_t ef_body(ann_ctx* ctx, int form, int with_evac, double* E, double* F, double* V) {
    hipError_t e;
    if ((F || V) && ctx->tune_fused == 1 && fused_supported(ctx, form)) {
        if ((e = symm_fused(ctx, ctx->da.inpb_net, ctx->d_Eatom, ctx->d_P)) != hipSuccess) return e;
        return force_reduce(ctx, ctx->d_P, ctx->d_Eatom, with_evac, F, E, V);
    }
    // the network's species tiles before G: their host set-up overlaps the G kernel
    if (form != ANN_FORM_PLAIN && (e = species_perm(ctx)) != hipSuccess) return e;
    if ((e = network_rows(ctx, form, ctx->d_G)) != hipSuccess) return e;
    if ((e = net_fwd(ctx, form, ctx->d_G, ctx->d_Eatom, F || V ? ctx->d_dEdG : nullptr, true)) != hipSuccess) return e;
    if (F || V) {
        if ((e = symm_bwd(ctx, form, ctx->d_dEdG, ctx->d_P)) != hipSuccess) return e;
        return force_reduce(ctx, ctx->d_P, ctx->d_Eatom, with_evac, F, E, V);
    }
    if (E) return struc_sums(ctx, ctx->d_Eatom, with_evac, nullptr, nullptr, E, nullptr, nullptr);
    return hipSuccess;
}

extern "C" int ann_energy_forces(ann_ctx* ctx, int form, int with_evac, double* E, double* F, double* V) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    RET_HIP(ensure_work(ctx), "allocation");
    ctx->G_valid = false;         // the call rewrites the context's G rows
    if (!(ctx->ef_graph_on && ctx->ef_warm)) {
        RET_HIP(ef_body(ctx, form, with_evac, E, F, V), "energy_forces");
        ctx->ef_warm = true;
        return ANN_OK;
    }
    // everything the captured launches bake in: arguments, buffers, grid and LDS
    // sizes, the species tiles of the network (set up on the host, so they are
    // brought up to date before the key is taken) and the FP64-evaluated inputs
    if (form != ANN_FORM_PLAIN) RET_HIP(species_perm(ctx), "species permutation");
    if (ctx->tune_queue == 1) RET_HIP(ensure_items(ctx), "work items");
    const NList& nl = ctx->nl;
    std::vector<uint64_t> key = {
        (uint64_t)form, (uint64_t)with_evac, (uint64_t)(uintptr_t)E, (uint64_t)(uintptr_t)F, (uint64_t)(uintptr_t)V,
        (uint64_t)(uintptr_t)nl.off, (uint64_t)(uintptr_t)nl.idx, (uint64_t)(uintptr_t)nl.img,
        (uint64_t)(uintptr_t)nl.rev, (uint64_t)(uintptr_t)nl.geo, (uint64_t)(uintptr_t)nl.nsp, (uint64_t)nl.maxn,
        (uint64_t)ctx->natoms, (uint64_t)ctx->nstruc, (uint64_t)(uintptr_t)ctx->d_pos,
        (uint64_t)(uintptr_t)ctx->d_species, (uint64_t)(uintptr_t)ctx->d_G, (uint64_t)(uintptr_t)ctx->d_P,
        (uint64_t)(uintptr_t)ctx->d_dEdG, (uint64_t)(uintptr_t)ctx->d_Eatom, (uint64_t)(uintptr_t)ctx->d_atomvir,
        (uint64_t)(uintptr_t)ctx->d_perm,
        (uint64_t)(uintptr_t)ctx->d_T, (uint64_t)(uintptr_t)ctx->d_chunks, (uint64_t)ctx->nchunks,
        (uint64_t)ctx->tune_apb[0], (uint64_t)ctx->tune_apb[1], (uint64_t)ctx->tune_apb[2],
        (uint64_t)ctx->tune_angular, (uint64_t)ctx->tune_queue, (uint64_t)ctx->tune_net, (uint64_t)ctx->tune_fused,
        (uint64_t)ctx->wsplit_valid, (uint64_t)ctx->wchunk_valid, ctx->batch_gen,
        (uint64_t)ctx->ntiles,
        (uint64_t)(uintptr_t)ctx->d_tiles, ctx->scaling_gen, (uint64_t)ctx->npre};
    if (!ctx->ef_exec || key != ctx->ef_key) {
        if (ctx->ef_exec) { hipGraphExecDestroy(ctx->ef_exec); ctx->ef_exec = nullptr; }
        hipStream_t user = ctx->stream;
        RET_HIP(hipEventRecord(ctx->gev[0], user), "graph event");
        RET_HIP(hipStreamWaitEvent(ctx->gstream, ctx->gev[0], 0), "graph wait");
        ctx->stream = ctx->gstream;
        ctx->capturing = true;
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->gstream, hipStreamCaptureModeRelaxed);
        hipError_t e2 = e == hipSuccess ? ef_body(ctx, form, with_evac, E, F, V) : e;
        hipError_t e3 = hipStreamEndCapture(ctx->gstream, &graph);
        ctx->capturing = false;
        ctx->stream = user;
        if (e != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
            return fail(ANN_ERR_HIP, std::string("energy/forces capture: ") +
                                         hipGetErrorString(e != hipSuccess ? e : (e2 != hipSuccess ? e2 : e3)));
        e = hipGraphInstantiate(&ctx->ef_exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) { ctx->ef_exec = nullptr; return fail(ANN_ERR_HIP, std::string("graph: ") + hipGetErrorString(e)); }
        ctx->ef_key = key;
    } else {
        RET_HIP(hipEventRecord(ctx->gev[0], ctx->stream), "graph event");
        RET_HIP(hipStreamWaitEvent(ctx->gstream, ctx->gev[0], 0), "graph wait");
    }
    RET_HIP(hipGraphLaunch(ctx->ef_exec, ctx->gstream), "graph launch");
    RET_HIP(hipEventRecord(ctx->gev[1], ctx->gstream), "graph event");
    RET_HIP(hipStreamWaitEvent(ctx->stream, ctx->gev[1], 0), "graph wait");
    return ANN_OK;
}

extern "C" int ann_ef_graph(ann_ctx* ctx, int enable) {
    CHECK_CTX(ctx);
    if (enable && !ctx->gstream) {
        ANN_HIP(hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
        ANN_HIP(hipEventCreateWithFlags(&ctx->gev[0], hipEventDisableTiming));
        ANN_HIP(hipEventCreateWithFlags(&ctx->gev[1], hipEventDisableTiming));
    }
    if (!enable && ctx->ef_exec) {
        ANN_HIP(hipStreamSynchronize(ctx->gstream));
        ANN_HIP(hipGraphExecDestroy(ctx->ef_exec));
        ctx->ef_exec = nullptr;
    }
    ctx->ef_graph_on = enable != 0;
    return ANN_OK;
}

extern "C" int ann_md_step(ann_ctx* ctx, int form, double dt, const double* mass, double* vel, double* F,
                           double* E, double* KE, double* V, int* moved) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!mass || !vel || !F) return fail(ANN_ERR_ARG, "null mass, velocity or force array");
    if (!(dt > 0.0)) return fail(ANN_ERR_ARG, "time step must be positive");
    RET_HIP(ensure_work(ctx), "allocation");
    const double skin = std::max(ctx->nl_rc - (double)ctx->hm.rc, 0.0);
    double* pos = const_cast<double*>(ctx->d_pos);
    if (ctx->natoms > ctx->md_cap) {
        ANN_HIP(dalloc(ctx->d_pos_ref, 3 * (size_t)ctx->natoms));
        ANN_HIP(dalloc(ctx->d_keatom, (size_t)ctx->natoms));
        if (!ctx->d_md_flag) ANN_HIP(hipMalloc(&ctx->d_md_flag, sizeof(int)));
        ctx->md_cap = ctx->natoms;
        ctx->pos_ref_valid = false;
    }
    if (!ctx->pos_ref_valid) {   // first step after a build: positions are still those of the build
        ANN_HIP(hipMemcpyAsync(ctx->d_pos_ref, pos, sizeof(double) * 3 * ctx->natoms, hipMemcpyDeviceToDevice,
                               ctx->stream));
        ANN_HIP(hipMemsetAsync(ctx->d_md_flag, 0, sizeof(int), ctx->stream));
        ctx->pos_ref_valid = true;
    }
    RET_HIP(md_kick_drift(ctx, dt, mass, F, vel, pos, 0.25 * skin * skin), "md_kick_drift");
    RET_HIP(refresh_geometry(ctx), "refresh_geometry");
    ctx->G_valid = false;
    ctx->gen++;
    RET_HIP(network_rows(ctx, form, ctx->d_G), "symm_fwd");
    RET_HIP(net_fwd(ctx, form, ctx->d_G, ctx->d_Eatom, ctx->d_dEdG, true), "net_fwd");
    RET_HIP(symm_bwd(ctx, form, ctx->d_dEdG, ctx->d_P), "symm_bwd");
    RET_HIP(force_reduce(ctx, ctx->d_P, ctx->d_Eatom, 1, F, E, V), "force_reduce");
    RET_HIP(md_kick(ctx, dt, mass, F, vel), "md_kick");
    if (KE) RET_HIP(struc_sums(ctx, nullptr, 0, nullptr, ctx->d_keatom, nullptr, nullptr, KE), "kinetic energy");
    if (moved) {
        RET_HIP(readback(ctx->d_md_flag, 1, ctx->d_rb + 8, ctx->stream), "flag");
        RET_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
        *moved = ((volatile const int*)ctx->h_rb)[8];
    }
    return ANN_OK;
}

extern "C" int ann_symm_tangent(ann_ctx* ctx, int form, const float* u, float* v) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!u || !v) return fail(ANN_ERR_ARG, "null argument");
    RET_HIP(symm_tangent(ctx, form, u, v), "symm_tangent");
    return ANN_OK;
}

extern "C" int ann_net_bwd2(ann_ctx* ctx, const float* G, const float* v, const double* gE, double* grad) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if (!G || !gE || !grad) return fail(ANN_ERR_ARG, "null argument");
    RET_HIP(net_bwd2(ctx, G, v, gE, grad), "net_bwd2");
    return ANN_OK;
}

extern "C" int ann_train_step(ann_ctx* ctx, int form, const double* Eref, const double* Fref, double wE, double wF,
                              double gamma, double* loss, double* grad) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    if (!Eref || !Fref) return fail(ANN_ERR_ARG, "null reference energies/forces");
    RET_HIP(ensure_work(ctx), "allocation");
    // graph replay: same inputs, outputs and weights as the captured step, and
    // no change of positions, list or batch since (every kernel reads its
    // buffers by pointer, and the Adam step counter lives on the device)
    const void* ptr[4] = {Eref, Fref, loss, grad};
    const bool same = ctx->gexec && ctx->gkey_gen == ctx->gen && ctx->gkey_form == form &&
                      std::memcmp(ptr, ctx->gkey_ptr, sizeof(ptr)) == 0 && ctx->gkey_w[0] == wE &&
                      ctx->gkey_w[1] == wF && ctx->gkey_w[2] == gamma;
    if (ctx->graph_on && ctx->train_warm && ctx->G_valid) {
        if (!same) {
            if (ctx->tune_queue == 1) RET_HIP(ensure_items(ctx), "work items");
            if (ctx->gexec) { hipGraphExecDestroy(ctx->gexec); ctx->gexec = nullptr; }
            hipStream_t user = ctx->stream;
            RET_HIP(hipEventRecord(ctx->gev[0], user), "graph event");
            RET_HIP(hipStreamWaitEvent(ctx->gstream, ctx->gev[0], 0), "graph wait");
            ctx->stream = ctx->gstream;
            ctx->capturing = true;
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamBeginCapture(ctx->gstream, hipStreamCaptureModeRelaxed);
            hipError_t e2 = e == hipSuccess ? train_step_impl(ctx, form, Eref, Fref, wE, wF, gamma, loss, grad) : e;
            hipError_t e3 = hipStreamEndCapture(ctx->gstream, &graph);
            ctx->capturing = false;
            ctx->stream = user;
            if (e != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
                return fail(ANN_ERR_HIP, std::string("train step capture: ") +
                                             hipGetErrorString(e != hipSuccess ? e : (e2 != hipSuccess ? e2 : e3)));
            e = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (e != hipSuccess) { ctx->gexec = nullptr; return fail(ANN_ERR_HIP, std::string("graph: ") + hipGetErrorString(e)); }
            ctx->gkey_gen = ctx->gen;
            ctx->gkey_form = form;
            std::memcpy(ctx->gkey_ptr, ptr, sizeof(ptr));
            ctx->gkey_w[0] = wE; ctx->gkey_w[1] = wF; ctx->gkey_w[2] = gamma;
        } else {
            RET_HIP(hipEventRecord(ctx->gev[0], ctx->stream), "graph event");
            RET_HIP(hipStreamWaitEvent(ctx->gstream, ctx->gev[0], 0), "graph wait");
        }
        RET_HIP(hipGraphLaunch(ctx->gexec, ctx->gstream), "graph launch");
        RET_HIP(hipEventRecord(ctx->gev[1], ctx->gstream), "graph event");
        RET_HIP(hipStreamWaitEvent(ctx->stream, ctx->gev[1], 0), "graph wait");
        return ANN_OK;
    }
    RET_HIP(train_step_impl(ctx, form, Eref, Fref, wE, wF, gamma, loss, grad), "train_step");
    ctx->train_warm = true;
    return ANN_OK;
}

extern "C" int ann_set_loss(ann_ctx* ctx, int loss, double delta, int norm) {
    CHECK_CTX(ctx);
    if (loss < ANN_LOSS_MSE || loss > ANN_LOSS_ASINH) return fail(ANN_ERR_ARG, "invalid loss");
    if (norm < ANN_NORM_IDENTITY || norm > ANN_NORM_LOG) return fail(ANN_ERR_ARG, "invalid normalization");
    if (!(delta > 0) || std::isinf(delta)) return fail(ANN_ERR_ARG, "loss width delta must be positive");
    ctx->loss_type = loss;
    ctx->loss_delta = delta;
    ctx->loss_norm = norm;
    ctx->gen++;                   // a captured training step bakes the loss in: capture again
    return ANN_OK;
}

extern "C" int ann_set_struc_weights(ann_ctx* ctx, const double* wE, const double* wF) {
    CHECK_CTX(ctx);
    ctx->d_wEs = wE;
    ctx->d_wFs = wF;
    ctx->gen++;
    return ANN_OK;
}

extern "C" int ann_train_graph(ann_ctx* ctx, int enable) {
    CHECK_CTX(ctx);
    if (enable && !ctx->gstream) {
        ANN_HIP(hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
        ANN_HIP(hipEventCreateWithFlags(&ctx->gev[0], hipEventDisableTiming));
        ANN_HIP(hipEventCreateWithFlags(&ctx->gev[1], hipEventDisableTiming));
    }
    if (!enable && ctx->gexec) {
        ANN_HIP(hipStreamSynchronize(ctx->gstream));
        ANN_HIP(hipGraphExecDestroy(ctx->gexec));
        ctx->gexec = nullptr;
    }
    ctx->graph_on = enable != 0;
    return ANN_OK;
}

// ---- autotuning ----
static double time_ms(ann_ctx* ctx, hipError_t (*f)(ann_ctx*, int), int arg, int reps, hipError_t& err) {
    hipEvent_t a = nullptr, b = nullptr;
    float ms = 0.0f;
    err = hipEventCreate(&a);
    if (err == hipSuccess) err = hipEventCreate(&b);
    if (err == hipSuccess) err = f(ctx, arg);                       // warm-up
    double best = 1e30;
    for (int trial = 0; trial < 3 && err == hipSuccess; ++trial) {  // the fastest of three timings
        err = hipEventRecord(a, ctx->stream);
        for (int r = 0; r < reps && err == hipSuccess; ++r) err = f(ctx, arg);
        if (err == hipSuccess) err = hipEventRecord(b, ctx->stream);
        if (err == hipSuccess) err = hipEventSynchronize(b);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, a, b);
        if (err == hipSuccess) best = std::min(best, (double)ms / reps);
    }
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    return best;
}
static hipError_t run_fwd(ann_ctx* c, int form) { return symm_fwd(c, form, c->d_G); }
static hipError_t run_bwd(ann_ctx* c, int form) { return symm_bwd(c, form, c->d_dEdG, c->d_P); }
static hipError_t run_tan(ann_ctx* c, int form) { return symm_tangent(c, form, c->d_u, c->d_v); }
static hipError_t run_ef(ann_ctx* c, int form) { return ef_body(c, form, 1, nullptr, c->d_F, nullptr); }
static hipError_t run_nb(ann_ctx* c, int) {
    c->grid_valid = false;      // the blocking policy is part of the cached grid
    return neighbors_build(c, c->nl_rc);
}

extern "C" int ann_autotune(ann_ctx* ctx, int form) {
    int rc = need_nlist(ctx); if (rc) return rc;
    if ((rc = check_form(form))) return rc;
    RET_HIP(ensure_work(ctx), "allocation");
    hipError_t e;
    ctx->tune_fused = 0;          // the E+F timings below run the separate kernels until the fused comparison
    RET_HIP(symm_fwd(ctx, form, ctx->d_G), "symm_fwd");
    RET_HIP(net_fwd(ctx, form, ctx->d_G, ctx->d_Eatom, ctx->d_dEdG), "net_fwd");
    RET_HIP(hipMemsetAsync(ctx->d_u, 0, sizeof(float) * 3 * (size_t)ctx->natoms, ctx->stream), "memset");
    const DevModel& M = ctx->hm;
    if (form == ANN_FORM_FAST && M.all_grid && M.moment_z > 0) {
        hipError_t (*runs[3])(ann_ctx*, int) = {run_fwd, run_bwd, run_tan};
        double best_total = 0.0;
        ctx->tune_angular = 0;
        for (int mode = 0; mode < 3; ++mode) {
            double best = 1e30;
            int pick = 1;
            for (int apb : {1, 2, 4}) {
                ctx->tune_apb[mode] = apb;
                const double t = time_ms(ctx, runs[mode], form, 3, e);
                if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
                if (t < best) { best = t; pick = apb; }
            }
            ctx->tune_apb[mode] = pick;
            if (mode < 2) best_total += best;
        }
        // persistent work queues against grid dispatch, timed as whole E+F steps
        // (the dispatch interacts with the stored moments and the network's
        // timing, which kernel-by-kernel timings miss)
        {
            ctx->tune_queue = 0;
            const double tg = time_ms(ctx, run_ef, form, 3, e);
            ctx->tune_queue = 1;
            const double tq = e == hipSuccess ? time_ms(ctx, run_ef, form, 3, e) : 0.0;
            if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
            ctx->tune_queue = tq < tg ? 1 : 0;
            RET_HIP(symm_fwd(ctx, form, ctx->d_G), "symm_fwd");
            best_total = time_ms(ctx, run_fwd, form, 3, e) + time_ms(ctx, run_bwd, form, 3, e);
            if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        }
        // moment form against the triplet tiles (G + adjoint)
        ctx->tune_angular = 1;
        const double tt = time_ms(ctx, run_fwd, form, 3, e) + time_ms(ctx, run_bwd, form, 3, e);
        if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        ctx->tune_angular = tt < best_total ? 1 : 0;
    }
    // network kernel (TRIPLET, FAST): vector-ALU tiles against the matrix cores,
    // timed as whole E+F steps with the separate kernels (chosen before the
    // fused comparison, which then times the separate path with its network)
    ctx->tune_fused = 0;
    if (form != ANN_FORM_PLAIN && vec_net_supported(ctx) && mfma_net_supported(M)) {
        double t[2];
        for (int q = 0; q < 2; ++q) {
            ctx->tune_net = q;
            t[q] = time_ms(ctx, run_ef, form, 3, e);
            if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        }
        ctx->tune_net = t[1] < t[0] ? 1 : 0;
    }
    // fused E+F (FAST, moment form) against the separate kernels, whole E+F steps
    if (fused_supported(ctx, form)) {
        double t[2];
        for (int q = 0; q < 2; ++q) {
            ctx->tune_fused = q;
            t[q] = time_ms(ctx, run_ef, form, 3, e);
            if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        }
        ctx->tune_fused = t[1] < t[0] ? 1 : 0;
    }
    // neighbour search blocking
    {
        double t[2];
        for (int sb = 0; sb < 2; ++sb) {
            ctx->tune_sb = sb;
            t[sb] = time_ms(ctx, run_nb, 0, 2, e);
            if (e != hipSuccess) return fail(ANN_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        }
        ctx->tune_sb = t[1] < t[0] ? 1 : 0;
        ctx->grid_valid = false;
        RET_HIP(neighbors_build(ctx, ctx->nl_rc), "neighbour build");
    }
    ctx->G_valid = false;
    ctx->pos_ref_valid = false;
    ctx->gen++;
    return ANN_OK;
}

extern "C" int ann_get_tuning(ann_ctx* ctx, int* t, int n) {
    CHECK_CTX(ctx);
    if (!t || n < 5) return fail(ANN_ERR_ARG, "tuning needs 5 entries");
    t[0] = ctx->tune_apb[0]; t[1] = ctx->tune_apb[1]; t[2] = ctx->tune_apb[2];
    t[3] = ctx->tune_angular; t[4] = ctx->tune_sb;
    if (n > 5) t[5] = ctx->tune_queue;
    if (n > 6) t[6] = ctx->tune_net;
    if (n > 7) t[7] = ctx->tune_fused;
    return ANN_OK;
}

extern "C" int ann_set_tuning(ann_ctx* ctx, const int* t, int n) {
    CHECK_CTX(ctx);
    if (!t || n < 5) return fail(ANN_ERR_ARG, "tuning needs 5 entries");
    for (int q = 0; q < 3; ++q)
        if (t[q] != 1 && t[q] != 2 && t[q] != 4) return fail(ANN_ERR_ARG, "atoms per workgroup must be 1, 2 or 4");
    if ((t[3] != 0 && t[3] != 1) || (t[4] != 0 && t[4] != 1) || (n > 5 && t[5] != 0 && t[5] != 1) ||
        (n > 6 && t[6] != 0 && t[6] != 1) || (n > 7 && t[7] != 0 && t[7] != 1))
        return fail(ANN_ERR_ARG, "invalid tuning switch");
    for (int q = 0; q < 3; ++q) ctx->tune_apb[q] = t[q];
    ctx->tune_angular = t[3];
    if (ctx->tune_sb != t[4]) ctx->grid_valid = false;
    ctx->tune_sb = t[4];
    if (n > 5) ctx->tune_queue = t[5];
    if (n > 6) ctx->tune_net = t[6];
    if (n > 7) ctx->tune_fused = t[7];
    ctx->gen++;
    return ANN_OK;
}

extern "C" int ann_profile(ann_ctx* ctx, int enable) {
    CHECK_CTX(ctx);
    int rc = drain_timers(ctx);
    if (rc) return rc;
    ctx->prof = enable != 0;
    ctx->timers.clear();
    return ANN_OK;
}

extern "C" int ann_profile_only(ann_ctx* ctx, const char* name) {
    CHECK_CTX(ctx);
    int rc = drain_timers(ctx);
    if (rc) return rc;
    ctx->prof_only = name ? name : "";
    return ANN_OK;
}

extern "C" int ann_profile_get(ann_ctx* ctx, const char* name, double* total_ms, int* count) {
    CHECK_CTX(ctx);
    int rc = drain_timers(ctx);
    if (rc) return rc;
    auto it = ctx->timers.find(name ? name : "");
    if (total_ms) *total_ms = it == ctx->timers.end() ? 0.0 : it->second.total_ms;
    if (count) *count = it == ctx->timers.end() ? 0 : it->second.count;
    return ANN_OK;
}
