// Aggregation multigrid hierarchy (host-side set-up); see mhs_amg.h.
#include "mhs_amg.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>

namespace mhs {

namespace {

void children(AmgLevel &C, int nf) {
    C.child_ptr.assign(C.n + 1, 0);
    for (int i = 0; i < nf; ++i)
        if (C.agg[i] >= 0) C.child_ptr[C.agg[i] + 1]++;
    for (int I = 0; I < C.n; ++I) C.child_ptr[I + 1] += C.child_ptr[I];
    C.child.assign(C.child_ptr[C.n], 0);
    std::vector<int> fill(C.child_ptr.begin(), C.child_ptr.end() - 1);
    for (int i = 0; i < nf; ++i)
        if (C.agg[i] >= 0) C.child[fill[C.agg[i]]++] = i;
}

// Sparsity of P'AP and the gather lists, coarse row by coarse row: the
// entries (agg(j), pos) of the coarse row's children are sorted locally.
// row_entries(i, emit) must call emit(j, pos) for every stored entry of
// finer row i (COARSE: j is already the coarse column, the entries being
// those of A P).  Requires children() to have run.
template <bool COARSE = false, class RowEntries>
void galerkin_pattern(AmgLevel &C, RowEntries row_entries) {
    C.row_ptr.assign(C.n + 1, 0);
    C.col.clear();
    C.src_ptr.assign(1, 0);
    C.src.clear();
    C.diag.assign(C.n, -1);
    std::vector<std::pair<int, int>> buf;
    for (int I = 0; I < C.n; ++I) {
        buf.clear();
        for (int k = C.child_ptr[I]; k < C.child_ptr[I + 1]; ++k)
            row_entries(C.child[k], [&](int j, int pos) {
                const int J = COARSE ? j : C.agg[j];
                if (J >= 0) buf.emplace_back(J, pos);
            });
        std::sort(buf.begin(), buf.end());
        for (size_t k = 0; k < buf.size(); ++k) {
            if (k == 0 || buf[k].first != buf[k - 1].first) {
                if (k > 0) C.src_ptr.push_back(int(C.src.size()));
                if (buf[k].first == I) C.diag[I] = int(C.col.size());
                C.col.push_back(buf[k].first);
            }
            C.src.push_back(buf[k].second);
        }
        if (!buf.empty()) C.src_ptr.push_back(int(C.src.size()));
        C.row_ptr[I + 1] = int(C.col.size());
    }
}

void build_ap(ApPattern &P, const Topology &T, const std::vector<int> &agg) {
    const int ns = T.nslices;
    std::vector<int> cnt(T.nfp, 0);
    std::vector<std::pair<int, long long>> buf;
    auto row = [&](int f) {
        buf.clear();
        const long long base = T.slice_ptr[f / kSlice] + f % kSlice;
        for (int e = 0; e < T.row_len[f]; ++e) {
            const long long pos = base + (long long)e * kSlice;
            const int J = agg[T.col[pos]];
            if (J >= 0) buf.emplace_back(J, pos);
        }
        std::sort(buf.begin(), buf.end());
    };
    for (int f = 0; f < T.nfp; ++f) {
        row(f);
        int n = 0;
        for (size_t k = 0; k < buf.size(); ++k) n += (k == 0 || buf[k].first != buf[k - 1].first);
        cnt[f] = n;
    }
    P.slice_ptr.assign(ns + 1, 0);
    for (int sl = 0; sl < ns; ++sl) {
        int w = 0;
        for (int l = 0; l < kSlice; ++l) w = std::max(w, cnt[sl * kSlice + l]);
        P.slice_ptr[sl + 1] = P.slice_ptr[sl] + (long long)w * kSlice;
    }
    const long long slots = P.slice_ptr[ns];
    P.col.assign(slots, 0);
    std::vector<int> len(slots, 0);
    // slot sizes, then the gather lists in slot order
    for (int f = 0; f < T.nfp; ++f) {
        row(f);
        const long long base = P.slice_ptr[f / kSlice] + f % kSlice;
        int e = -1;
        for (size_t k = 0; k < buf.size(); ++k) {
            if (k == 0 || buf[k].first != buf[k - 1].first) {
                ++e;
                P.col[base + (long long)e * kSlice] = buf[k].first;
            }
            len[base + (long long)e * kSlice]++;
        }
    }
    P.src_ptr.assign(slots + 1, 0);
    for (long long q = 0; q < slots; ++q) P.src_ptr[q + 1] = P.src_ptr[q] + len[q];
    P.src.assign(P.src_ptr[slots], 0);
    std::vector<int> fill(P.src_ptr.begin(), P.src_ptr.end() - 1);
    for (int f = 0; f < T.nfp; ++f) {
        row(f);
        const long long base = P.slice_ptr[f / kSlice] + f % kSlice;
        int e = -1;
        for (size_t k = 0; k < buf.size(); ++k) {
            if (k == 0 || buf[k].first != buf[k - 1].first) ++e;
            P.src[fill[base + (long long)e * kSlice]++] = int(buf[k].second);
        }
    }
}

// Smoothed-aggregation transfer F -> C (C.agg set, F's CSR pattern known):
// the patterns of P, P', T = A_F P and A_C = P' T, with the gather / pair
// lists that let every solve fill their values on the device.
void sa_transfer(AmgLevel &F, AmgLevel &C) {
    const int n = F.n;
    F.sa_out = true;
    F.p_ptr.assign(n + 1, 0);
    F.p_col.clear();
    F.p_own.assign(n, -1);
    F.p_src_ptr.assign(1, 0);
    F.p_src.clear();
    std::vector<std::pair<int, int>> buf;
    for (int i = 0; i < n; ++i) {
        buf.clear();
        for (int k = F.row_ptr[i]; k < F.row_ptr[i + 1]; ++k) buf.emplace_back(C.agg[F.col[k]], k);
        std::sort(buf.begin(), buf.end());
        for (size_t k = 0; k < buf.size(); ++k) {
            if (k == 0 || buf[k].first != buf[k - 1].first) {
                if (k > 0) F.p_src_ptr.push_back(int(F.p_src.size()));
                if (buf[k].first == C.agg[i]) F.p_own[i] = int(F.p_col.size());
                F.p_col.push_back(buf[k].first);
            }
            F.p_src.push_back(buf[k].second);
        }
        if (!buf.empty()) F.p_src_ptr.push_back(int(F.p_src.size()));
        F.p_ptr[i + 1] = int(F.p_col.size());
    }
    // P' (rows of C)
    F.pt_ptr.assign(C.n + 1, 0);
    for (int J : F.p_col) F.pt_ptr[J + 1]++;
    for (int I = 0; I < C.n; ++I) F.pt_ptr[I + 1] += F.pt_ptr[I];
    F.pt_col.assign(F.p_col.size(), 0);
    F.pt_map.assign(F.p_col.size(), 0);
    {
        std::vector<int> fill(F.pt_ptr.begin(), F.pt_ptr.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int q = F.p_ptr[i]; q < F.p_ptr[i + 1]; ++q) {
                const int e = fill[F.p_col[q]]++;
                F.pt_col[e] = i;
                F.pt_map[e] = q;
            }
        F.p_tpos.assign(F.p_col.size(), 0);
        for (size_t e = 0; e < F.pt_map.size(); ++e) F.p_tpos[F.pt_map[e]] = int(e);
    }
    // T = A P
    std::vector<std::array<int, 3>> tb;
    F.t_ptr.assign(n + 1, 0);
    F.t_col.clear();
    F.t_pair_ptr.assign(1, 0);
    F.t_pa.clear();
    F.t_pb.clear();
    for (int i = 0; i < n; ++i) {
        tb.clear();
        for (int k = F.row_ptr[i]; k < F.row_ptr[i + 1]; ++k) {
            const int j = F.col[k];
            for (int q = F.p_ptr[j]; q < F.p_ptr[j + 1]; ++q) tb.push_back({F.p_col[q], k, q});
        }
        std::sort(tb.begin(), tb.end());
        for (size_t k = 0; k < tb.size(); ++k) {
            if (k == 0 || tb[k][0] != tb[k - 1][0]) {
                if (k > 0) F.t_pair_ptr.push_back(int(F.t_pa.size()));
                F.t_col.push_back(tb[k][0]);
            }
            F.t_pa.push_back(tb[k][1]);
            F.t_pb.push_back(tb[k][2]);
        }
        if (!tb.empty()) F.t_pair_ptr.push_back(int(F.t_pa.size()));
        F.t_ptr[i + 1] = int(F.t_col.size());
    }

    // A_C = P' T
    C.sa_in = true;
    C.row_ptr.assign(C.n + 1, 0);
    C.col.clear();
    C.src_ptr.assign(1, 0);
    C.src.clear();
    C.src2.clear();
    C.diag.assign(C.n, -1);
    for (int I = 0; I < C.n; ++I) {
        tb.clear();
        for (int e = F.pt_ptr[I]; e < F.pt_ptr[I + 1]; ++e) {
            const int i = F.pt_col[e], q = F.pt_map[e];
            for (int t = F.t_ptr[i]; t < F.t_ptr[i + 1]; ++t) tb.push_back({F.t_col[t], q, t});
        }
        std::sort(tb.begin(), tb.end());
        for (size_t k = 0; k < tb.size(); ++k) {
            if (k == 0 || tb[k][0] != tb[k - 1][0]) {
                if (k > 0) C.src_ptr.push_back(int(C.src.size()));
                if (tb[k][0] == I) C.diag[I] = int(C.col.size());
                C.col.push_back(tb[k][0]);
            }
            C.src.push_back(tb[k][1]);
            C.src2.push_back(tb[k][2]);
        }
        if (!tb.empty()) C.src_ptr.push_back(int(C.src.size()));
        C.row_ptr[I + 1] = int(C.col.size());
    }
}

// Block factors for one transfer: 2 along the coarsenable directions whose
// strength is at least theta times the strongest coarsenable one.
std::array<int, 3> choose_blocks(const Strength &s, double theta, const std::vector<int> &cur) {
    const int ninst = int(cur.size() / 3);
    bool can[3] = {false, false, false};
    for (int t = 0; t < ninst; ++t)
        for (int d = 0; d < 3; ++d) can[d] = can[d] || cur[3 * t + d] > 1;
    double smax = 0.0;
    for (int d = 0; d < 3; ++d)
        if (can[d]) smax = std::max(smax, s[d]);
    std::array<int, 3> b{1, 1, 1};
    for (int d = 0; d < 3; ++d)
        if (can[d] && (!(smax > 0.0) || s[d] >= theta * smax)) b[d] = 2;
    return b;
}

}  // namespace

std::string instance_dims(const Topology &T, const std::vector<int> &ijk, std::vector<int> &dims) {
    const int ninst = T.ninst;
    if ((int)ijk.size() != 3 * T.nc) return "logical cell indices: need 3 per cell";
    dims.assign(3 * ninst, 0);
    for (int t = 0; t < ninst; ++t)
        for (int c = T.inst_cellpos[t]; c < T.inst_cellpos[t + 1]; ++c)
            for (int d = 0; d < 3; ++d) {
                if (ijk[3 * c + d] < 0) return "logical cell indices must be non-negative";
                dims[3 * t + d] = std::max(dims[3 * t + d], ijk[3 * c + d] + 1);
            }
    return "";
}

std::vector<std::array<int, 3>> coarsening_plan(const Strength &strength, const std::vector<int> &dims) {
    const int ninst = int(dims.size() / 3);
    Strength s = strength;
    for (double &v : s)
        if (!(v > 0.0) || !std::isfinite(v)) v = 0.0;
    std::vector<int> cur = dims;
    std::vector<std::array<int, 3>> plan;
    for (;;) {
        if (!plan.empty()) {
            long long mx = 0;
            for (int t = 0; t < ninst; ++t)
                mx = std::max(mx, (long long)cur[3 * t] * cur[3 * t + 1] * cur[3 * t + 2]);
            if (mx <= kMaxCoarsest || plan.size() >= 40) break;
        }
        const std::array<int, 3> b = choose_blocks(s, plan.empty() ? 0.5 : 0.25, cur);
        plan.push_back(b);
        for (int t = 0; t < ninst; ++t)
            for (int d = 0; d < 3; ++d) cur[3 * t + d] = (cur[3 * t + d] + b[d] - 1) / b[d];
        const Strength s0 = s;
        for (int d = 0; d < 3; ++d) s[d] = s0[d] * b[(d + 1) % 3] * b[(d + 2) % 3] / b[d];
    }
    return plan;
}

std::string build_amg(AmgHierarchy &H, const Topology &T, const std::vector<int> &ijk, const Strength &strength,
                      bool smoothed) {
    H = AmgHierarchy();
    const int ninst = T.ninst;
    if (T.slice_ptr[T.nslices] > 0x7fffffffLL)
        return "face matrix too large for the 32-bit multigrid gather lists";
    std::vector<int> dims;
    std::string ed = instance_dims(T, ijk, dims);
    if (!ed.empty()) return ed;
    H.blocks = coarsening_plan(strength, dims);
    H.lev.emplace_back();  // level 0 is the SELL face matrix
    // ---- level 1: faces -> blocks of their owner cells
    {
        const std::array<int, 3> &b0 = H.blocks[0];
        AmgLevel L1;
        L1.inst_ptr.assign(ninst + 1, 0);
        L1.dims.resize(3 * ninst);
        for (int t = 0; t < ninst; ++t) {
            long long m = 1;
            for (int d = 0; d < 3; ++d) {
                L1.dims[3 * t + d] = (dims[3 * t + d] + b0[d] - 1) / b0[d];
                m *= L1.dims[3 * t + d];
            }
            L1.inst_ptr[t + 1] = L1.inst_ptr[t] + int(m);
        }
        L1.n = L1.inst_ptr[ninst];
        L1.agg.assign(T.nfp, -1);
        std::vector<int> inst_of_cell(T.nc);
        for (int t = 0; t < ninst; ++t)
            for (int c = T.inst_cellpos[t]; c < T.inst_cellpos[t + 1]; ++c) inst_of_cell[c] = t;
        for (int f = 0; f < T.nfp; ++f) {
            const int h = T.dof_hf[f];
            if (h < 0) continue;
            const int c = T.hf_cell[h], t = inst_of_cell[c];
            const int *bd = &L1.dims[3 * t];
            const int I = ijk[3 * c] / b0[0], J = ijk[3 * c + 1] / b0[1], K = ijk[3 * c + 2] / b0[2];
            L1.agg[f] = L1.inst_ptr[t] + I + bd[0] * (J + bd[1] * K);
        }
        children(L1, T.nfp);
        // A_1 = P' (A P): each entry sums the A P slots of its row's faces
        // (about 4 per face row) instead of the face matrix's entries (11)
        build_ap(H.ap, T, L1.agg);
        const ApPattern &ap = H.ap;
        galerkin_pattern<true>(L1, [&](int f, auto emit) {
            const long long base = ap.slice_ptr[f / kSlice] + f % kSlice;
            const int w = int((ap.slice_ptr[f / kSlice + 1] - ap.slice_ptr[f / kSlice]) / kSlice);
            for (int e = 0; e < w; ++e) {
                const long long q = base + (long long)e * kSlice;
                if (ap.src_ptr[q + 1] > ap.src_ptr[q]) emit(ap.col[q], int(q));   // padding slots sum nothing
            }
        });
        H.lev.push_back(std::move(L1));
    }
    // ---- coarser levels: group the blocks again until every instance is small
    bool sa_done = false;
    for (size_t l = 1; l < H.blocks.size(); ++l) {
        const AmgLevel &F = H.lev.back();
        const std::array<int, 3> &bl = H.blocks[l];
        AmgLevel C;
        C.inst_ptr.assign(ninst + 1, 0);
        C.dims.resize(3 * ninst);
        for (int t = 0; t < ninst; ++t) {
            long long m = 1;
            for (int d = 0; d < 3; ++d) {
                C.dims[3 * t + d] = (F.dims[3 * t + d] + bl[d] - 1) / bl[d];
                m *= C.dims[3 * t + d];
            }
            C.inst_ptr[t + 1] = C.inst_ptr[t] + int(m);
        }
        C.n = C.inst_ptr[ninst];
        C.agg.assign(F.n, -1);
        for (int t = 0; t < ninst; ++t) {
            const int *fd = &F.dims[3 * t], *cd = &C.dims[3 * t];
            for (int r = F.inst_ptr[t]; r < F.inst_ptr[t + 1]; ++r) {
                const int loc = r - F.inst_ptr[t];
                const int i = loc % fd[0], j = (loc / fd[0]) % fd[1], k = loc / (fd[0] * fd[1]);
                C.agg[r] = C.inst_ptr[t] + i / bl[0] + cd[0] * (j / bl[1] + cd[1] * (k / bl[2]));
            }
        }
        AmgLevel &Fr = H.lev.back();
        children(C, F.n);
        // smooth one transfer: the first from a level with at most an eighth
        // as many rows as there are faces (cheap next to the face level),
        // most of the gain in iterations, while the other levels keep the
        // sparse stencils of piecewise-constant aggregation (smoothing widens
        // every stencil)
        if (smoothed && !sa_done && 8LL * Fr.n <= (long long)T.nf_total) {
            sa_done = true;
            sa_transfer(Fr, C);
            // the up-sweep stores M = P - w D^-1 T on T's pattern, which holds
            // P's (A holds its diagonal; both rows sorted)
            Fr.t_p.assign(Fr.t_col.size(), -1);
            for (int i = 0; i < Fr.n; ++i) {
                int t = Fr.t_ptr[i];
                for (int q = Fr.p_ptr[i]; q < Fr.p_ptr[i + 1]; ++q) {
                    while (t < Fr.t_ptr[i + 1] && Fr.t_col[t] < Fr.p_col[q]) ++t;
                    if (t == Fr.t_ptr[i + 1] || Fr.t_col[t] != Fr.p_col[q])
                        return "smoothed transfer: a row of P outside the pattern of A P";
                    Fr.t_p[t] = q;
                }
            }
        } else {
            galerkin_pattern(C, [&](int r, auto emit) {
                for (int k = Fr.row_ptr[r]; k < Fr.row_ptr[r + 1]; ++k) emit(Fr.col[k], k);
            });
        }
        H.lev.push_back(std::move(C));
    }
    const AmgLevel &Lc = H.lev.back();
    for (int t = 0; t < ninst; ++t)
        if (Lc.inst_ptr[t + 1] - Lc.inst_ptr[t] > kMaxCoarsest) return "multigrid hierarchy does not coarsen";
    H.coarse_off.assign(ninst + 1, 0);
    for (int t = 0; t < ninst; ++t) {
        const int m = Lc.inst_ptr[t + 1] - Lc.inst_ptr[t];
        H.coarse_off[t + 1] = H.coarse_off[t] + m * m;
    }
    for (size_t l = 1; l < H.lev.size(); ++l)
        for (int I = 0; I < H.lev[l].n; ++I)
            if (H.lev[l].diag[I] < 0) return "coarse row without a diagonal entry";
    return "";
}

}  // namespace mhs
