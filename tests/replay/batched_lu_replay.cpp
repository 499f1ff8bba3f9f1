// Host replay of the batched LU kernel's address arithmetic
// (hygondoubleblockedfactorizations_amd/csrc/batched.hip: order_kernel,
// batched_kernel<kPrecFP32, true>, lu_matrix with its panel in LDS,
// tile_sweep<kPrecFP32, false, 1> with its two operand buffers in the panel
// array and, for trailing orders up to the small size class,
// tile_update_small<kPrecFP32, false> staging in the panel).
//
// Every phase between two workgroup barriers runs for tid = 0..255 in turn
// (or, where the phase's accesses do not depend on the thread split, as one
// loop over the same index range);
// the wave butterflies run lane by lane.  Every access to the packed matrix
// buffer, to ipiv and to the workgroup arrays goes through a checked
// accessor that counts accesses outside (a) the whole buffer, (b) the
// current matrix's lda x n block, (c) its n x n part.  Values are computed
// exactly as the kernel computes them (FP32), so the pivot sequence (which
// decides the addresses of the row exchanges) is the kernel's.
//
// Input (little endian): int32 batch, int64 total, int32 n[batch],
// int32 lda[batch], int64 offsets[batch], int64 ipiv_offsets[batch],
// float flat[total].  Output: float flat[total], int32 ipiv[sum max(n,1)],
// int32 info[batch].  Optional third argument: the small size class's
// largest trailing order (the library's default, 96, when absent).  Prints
// the count of out-of-range accesses; exit 1 if any.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr int kNB = 32, kT = 64, kMaxOrder = 4096, kLuMaxOrder = 1024, kBuckets = 256,
              kThreads = 256, kSmallT = 32;
int small_max = 96;

struct Checked {
    std::vector<float> flat;
    std::vector<int> ipiv;
    long long lo = 0, hi = 0, plo = 0, phi = 0;  // current matrix / ipiv range
    long long base = 0, lda = 0;
    int n = 0;
    long long violations = 0;
    float sink = 0.0f;
    int isink = 0;

    void report(const char* what, long long a, long long b)
    {
        if (violations++ < 20) std::printf("out of range: %s (%lld, %lld)\n", what, a, b);
    }
    float& A(long long r, long long c)
    {
        const long long i = base + r + c * lda;
        if (r < 0 || r >= n || c < 0 || c >= n) report("outside n x n", r, c);
        if (i < lo || i >= hi) report("outside lda x n block", r, c);
        if (i < 0 || i >= static_cast<long long>(flat.size())) {
            report("outside buffer", r, c);
            return sink;
        }
        return flat[i];
    }
    int& piv(long long j)
    {
        const long long i = plo + j;
        if (i < plo || i >= phi || i >= static_cast<long long>(ipiv.size())) {
            report("ipiv", j, 0);
            return isink;
        }
        return ipiv[i];
    }
};

template <typename T, int R, int C>
struct Lds {
    T v[R][C];
    long long* violations;
    T& at(int r, int c)
    {
        static T s;
        if (r < 0 || r >= R || c < 0 || c >= C) {
            ++*violations;
            return s;
        }
        return v[r][c];
    }
};

using Panel = Lds<float, kLuMaxOrder, kNB + 1>;

// C -= L R with C = A(r1.., r1..) m x nc, L = A(r1.., kL..) and R =
// A(kL.., r1..), K = kw (<= 64), tile by tile (64 x 64, down each tile
// column, then the next), each tile in 32-wide k slices; the operands of
// step s staged in buffer s % 2 of the panel array (k-major, 2 x kNB x 65
// floats per buffer); each tile's C block is read at its first slice and
// written after its last (the products run as C^T -= R^T L^T, so in the
// MFMA accumulator layout lane l holds row l % 32, columns 8 (r / 4) + 4
// (l / 32) + r % 4).
void tile_sweep(Checked& M, Panel& pnl, int r1, int kL, int m, int nc, int kw)
{
    const int tiles_r = (m + kT - 1) / kT, tiles_c = (nc + kT - 1) / kT;
    const int slices = (kw + kNB - 1) / kNB;
    constexpr int kTileFloats = 2 * kNB * (kT + 1);
    auto buf = [&](int t, int half, int kk, int i) -> float& {
        const int idx = (t & 1) * kTileFloats + half * kNB * (kT + 1) + kk * (kT + 1) + i;
        return pnl.at(idx / (kNB + 1), idx % (kNB + 1));
    };
    std::vector<float> acc(static_cast<size_t>(kThreads) * 16), cv(acc.size());
    int t = 0;
    for (int tj = 0; tj < tiles_c; ++tj)
        for (int ti = 0; ti < tiles_r; ++ti)
            for (int sl = 0; sl < slices; ++sl, ++t) {
                const int r0 = ti * kT, c0 = tj * kT, k0 = sl * kNB;
                for (int tid = 0; tid < kThreads; ++tid)
                    for (int q = 0; q < kNB * kT / kThreads; ++q) {
                        const int e = tid + kThreads * q;
                        const int kk = k0 + e / kT, i = e % kT;
                        float va = 0.0f, vb = 0.0f;
                        if (kk < kw) {
                            if (r0 + i < m) va = M.A(r1 + r0 + i, kL + kk);
                            if (c0 + i < nc) vb = M.A(kL + kk, r1 + c0 + i);
                        }
                        buf(t, 0, e / kT, i) = va;
                        buf(t, 1, e / kT, i) = vb;
                    }
                for (int tid = 0; tid < kThreads; ++tid) {
                    const int lane = tid & 63, wave = tid >> 6, wm = wave & 1, wn = wave >> 1;
                    const int row = wm * 32 + (lane & 31);
                    const int gr = r0 + row;
                    for (int r = 0; r < 16; ++r) {
                        const int lc = wn * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                        const bool ok = gr < m && c0 + lc < nc;
                        float& a = acc[tid * 16 + r];
                        if (sl == 0) {
                            cv[tid * 16 + r] = ok ? M.A(r1 + gr, r1 + c0 + lc) : 0.0f;
                            a = 0.0f;
                        }
                        for (int kk = 0; kk < kNB; ++kk) a = std::fma(buf(t, 1, kk, lc), buf(t, 0, kk, row), a);
                        if (sl == slices - 1 && ok) M.A(r1 + gr, r1 + c0 + lc) = cv[tid * 16 + r] - a;
                    }
                }
            }
}

// The small size class: wave w takes tiles w, w + 4, ... of the 32-blocks of
// C (column-major tile order), stages its operands one 32-wide k slice at a
// time in its 2 x 32 x 33 slice of the panel array and writes its 32 x 32
// tile after the last slice.
void tile_update_small(Checked& M, Panel& pnl, int r1, int kL, int m, int nc, int kw)
{
    const int tr = (m + kSmallT - 1) / kSmallT, tc = (nc + kSmallT - 1) / kSmallT, count = tr * tc;
    auto slot = [&](int wave, int half, int kk, int i) -> float& {
        const int idx = wave * 2 * kNB * (kSmallT + 1) + half * kNB * (kSmallT + 1)
                        + kk * (kSmallT + 1) + i;
        return pnl.at(idx / (kNB + 1), idx % (kNB + 1));
    };
    for (int wave = 0; wave < 4; ++wave)
        for (int t = wave; t < count; t += 4) {
            const int r0 = (t % tr) * kSmallT, c0 = (t / tr) * kSmallT;
            std::vector<float> acc(64 * 16, 0.0f);
            for (int k0 = 0; k0 < kw; k0 += kNB) {
                for (int lane = 0; lane < 64; ++lane)
                    for (int e = lane; e < kNB * kSmallT; e += 64) {
                        const int i = e % kSmallT, kk = e / kSmallT;
                        float va = 0.0f, vb = 0.0f;
                        if (k0 + kk < kw) {
                            if (r0 + i < m) va = M.A(r1 + r0 + i, kL + k0 + kk);
                            if (c0 + i < nc) vb = M.A(kL + k0 + kk, r1 + c0 + i);
                        }
                        slot(wave, 0, kk, i) = va;
                        slot(wave, 1, kk, i) = vb;
                    }
                for (int lane = 0; lane < 64; ++lane)
                    for (int r = 0; r < 16; ++r) {
                        const int lc = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                        for (int kk = 0; kk < kNB; ++kk)
                            acc[lane * 16 + r] = std::fma(slot(wave, 1, kk, lc), slot(wave, 0, kk, lane & 31),
                                                          acc[lane * 16 + r]);
                    }
            }
            for (int lane = 0; lane < 64; ++lane) {
                const int gr = r0 + (lane & 31);
                if (gr >= m) continue;
                for (int r = 0; r < 16; ++r) {
                    const int lc = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    if (c0 + lc < nc) M.A(r1 + gr, r1 + c0 + lc) -= acc[lane * 16 + r];
                }
            }
        }
}

void lu_matrix(Checked& M, int n, int* info)
{
    static Panel pnl;
    Lds<int, 1, kNB> piv_s;
    static Lds<int, 1, kLuMaxOrder> rowmap;
    Lds<int, 1, 2 * kNB> mv_src, mv_dst;
    rowmap.violations = mv_src.violations = mv_dst.violations = &M.violations;
    static Lds<int, 1, kLuMaxOrder> piv_all;
    piv_all.violations = &M.violations;
    Lds<float, 8, kNB + 1> cand;
    Lds<float, 2, 4> cand_v;
    Lds<int, 2, 4> cand_i;
    cand.violations = cand_v.violations = cand_i.violations = &M.violations;
    pnl.violations = piv_s.violations = &M.violations;
    static Lds<float, kNB, kNB + 1> lba;
    lba.violations = &M.violations;
    for (int k = 0; k < n; k += kNB) {
        const int kb = std::min(kNB, n - k);
        const int m = n - k;
        const bool second = (k / kNB) & 1;
        const int k0 = second ? k - kNB : k;
        // 1. panel -> LDS, rows tid, tid + 256, ... per thread
        for (int tid = 0; tid < kThreads; ++tid)
            for (int r = tid; r < m; r += kThreads)
                for (int c = 0; c < kNB; ++c) pnl.at(r, c) = c < kb ? M.A(k + r, k + c) : 0.0f;
        // each wave's arg-max candidate (the kernel's one wavefront max of
        // the key (|v| bits, ~index); no candidate: key 0) and a copy of its
        // row, per column parity
        auto publish = [&](int par, std::vector<float>& bv, std::vector<int>& bi) {
            for (int w = 0; w < 4; ++w) {
                unsigned long long kmax = 0;
                for (int l = 0; l < 64; ++l) {
                    const float v = bv[w * 64 + l];
                    const int ii = bi[w * 64 + l];
                    unsigned int bits;
                    std::memcpy(&bits, &v, sizeof(bits));
                    const unsigned long long key =
                        (ii < m && v >= 0.0f)
                            ? (static_cast<unsigned long long>(bits) << 32) | static_cast<unsigned int>(~ii)
                            : 0ULL;
                    kmax = std::max(kmax, key);
                }
                float wv = -1.0f;
                int wi = m;
                if (kmax != 0) {
                    const unsigned int hi = static_cast<unsigned int>(kmax >> 32);
                    std::memcpy(&wv, &hi, sizeof(wv));
                    wi = ~static_cast<int>(static_cast<unsigned int>(kmax));
                }
                for (int l = 0; l < 64; ++l) {
                    bv[w * 64 + l] = wv;
                    bi[w * 64 + l] = wi;
                }
                cand_v.at(par, w) = bv[w * 64];
                cand_i.at(par, w) = bi[w * 64];
                for (int l = 0; l < kNB; ++l)
                    cand.at(par * 4 + w, l) = bi[w * 64] < m ? pnl.at(bi[w * 64], l) : 0.0f;
            }
        };
        {
            std::vector<float> bv(kThreads, -1.0f);
            std::vector<int> bi(kThreads, m);
            for (int tid = 0; tid < kThreads; ++tid)
                for (int r = tid; r < m; r += kThreads) {
                    const float v = std::fabs(pnl.at(r, 0));
                    if (v > bv[tid]) {
                        bv[tid] = v;
                        bi[tid] = r;
                    }
                }
            publish(0, bv, bi);
        }
        for (int j = 0; j < kb; ++j) {
            const int par = j & 1;
            float pv = cand_v.at(par, 0);
            int p = cand_i.at(par, 0), w = 0;
            for (int q = 1; q < 4; ++q)
                if (cand_v.at(par, q) > pv || (cand_v.at(par, q) == pv && cand_i.at(par, q) < p)) {
                    pv = cand_v.at(par, q);
                    p = cand_i.at(par, q);
                    w = q;
                }
            const bool none = p < j || p >= m;
            if (none) p = j;
            auto prow = [&](int c) { return none ? pnl.at(j, c) : cand.at(par * 4 + w, c); };
            const float piv = prow(j);
            M.piv(k + j) = k + p + 1;
            piv_s.at(0, j) = p;
            piv_all.at(0, k + j) = k + p;
            if (piv == 0.0f && *info == 0) *info = k + j + 1;
            const bool scale = piv != 0.0f;
            const bool recip = std::fabs(piv) >= FLT_MIN;
            const float rinv = 1.0f / piv;
            // the exchange by the wave owning row p (lanes 0..31), before its
            // lane p % 64 updates row p; no other wave touches rows j, p
            if (p != j)
                for (int l = 0; l < kNB; ++l) {
                    const float t = pnl.at(j, l);
                    pnl.at(j, l) = prow(l);
                    pnl.at(p, l) = t;
                }
            std::vector<float> bv(kThreads, -1.0f);
            std::vector<int> bi(kThreads, m);
            for (int tid = 0; tid < kThreads; ++tid)
                for (int r = tid; r < m; r += kThreads) {
                    if (r <= j) continue;
                    if (scale) {
                        const float x = pnl.at(r, j);
                        const float l = recip ? x * rinv : x / piv;
                        pnl.at(r, j) = l;
                        for (int c = j + 1; c < kb; ++c) pnl.at(r, c) -= l * prow(c);
                    }
                    if (j + 1 < kb) {
                        const float v = std::fabs(pnl.at(r, j + 1));
                        if (v > bv[tid]) {
                            bv[tid] = v;
                            bi[tid] = r;
                        }
                    }
                }
            if (j + 1 < kb) publish(par ^ 1, bv, bi);
        }
        // 2. panel -> global; interchanges to the other columns
        for (int tid = 0; tid < kThreads; ++tid)
            for (int r = tid; r < m; r += kThreads)
                for (int c = 0; c < kb; ++c) M.A(k + r, k + c) = pnl.at(r, c);
        const int c1 = k + kb;
        const bool right = c1 < n;
        if (!right && !second) continue;
        // the panel's interchanges composed into a move list (rowmap: whose
        // content ends at each panel row)
        if (right && kb != kNB) M.report("right columns after a narrow panel", k, kb);
        for (int tid = 0; tid < kb; ++tid) {
            rowmap.at(0, tid) = tid;
            rowmap.at(0, piv_s.at(0, tid)) = piv_s.at(0, tid);
        }
        for (int jj = 0; jj < kb; ++jj) {
            const int pp = piv_s.at(0, jj);
            if (pp != jj) std::swap(rowmap.at(0, jj), rowmap.at(0, pp));
        }
        for (int tid = 0; tid < 2 * kNB; ++tid) {
            int dst = -1;
            if (tid < kb) {
                dst = tid;
            } else if (tid >= kNB && tid - kNB < kb) {
                const int u = tid - kNB;
                const int pp = piv_s.at(0, u);
                bool first = pp >= kb;
                for (int v = 0; v < u; ++v) first = first && piv_s.at(0, v) != pp;
                if (first) dst = pp;
            }
            const int src = dst >= 0 ? rowmap.at(0, dst) : dst;
            mv_dst.at(0, tid) = dst;
            mv_src.at(0, tid) = (dst >= 0 && src != dst) ? src : -1;
        }
        if (second) {
            // 2b. the moves on the first panel's columns; L_ba to LDS
            for (int tid = 0; tid < kNB; ++tid) {
                const int c = k0 + tid;
                float x[kNB], y[kNB];
                for (int ii = 0; ii < kNB; ++ii)
                    x[ii] = ii < kb ? M.A(k + (mv_src.at(0, ii) >= 0 ? mv_src.at(0, ii) : ii), c) : 0.0f;
                for (int u = 0; u < kNB; ++u)
                    y[u] = mv_src.at(0, kNB + u) >= 0 ? M.A(k + mv_src.at(0, kNB + u), c) : 0.0f;
                for (int u = 0; u < kNB; ++u)
                    if (mv_src.at(0, kNB + u) >= 0) M.A(k + mv_dst.at(0, kNB + u), c) = y[u];
                for (int ii = 0; ii < kNB; ++ii) {
                    if (mv_src.at(0, ii) >= 0) M.A(k + ii, c) = x[ii];
                    lba.at(ii, tid) = x[ii];
                }
            }
        }
        if (right) {
            // 3. each column right of the panel: the moves (every read
            //    before any write), (second panel) minus L_ba times the first
            //    panel's U rows, then U := L11^{-1} (.) in registers
            for (int c = c1; c < n; ++c) {
                float x[kNB], y[kNB];
                for (int ii = 0; ii < kNB; ++ii)
                    x[ii] = M.A(k + (mv_src.at(0, ii) >= 0 ? mv_src.at(0, ii) : ii), c);
                for (int u = 0; u < kNB; ++u)
                    y[u] = mv_src.at(0, kNB + u) >= 0 ? M.A(k + mv_src.at(0, kNB + u), c) : 0.0f;
                for (int u = 0; u < kNB; ++u)
                    if (mv_src.at(0, kNB + u) >= 0) M.A(k + mv_dst.at(0, kNB + u), c) = y[u];
                if (second) {
                    float xa[kNB];
                    for (int q = 0; q < kNB; ++q) xa[q] = M.A(k0 + q, c);
                    for (int ii = 0; ii < kNB; ++ii) {
                        float v = x[ii];
                        for (int q = 0; q < kNB; ++q) v -= lba.at(ii, q) * xa[q];
                        x[ii] = v;
                    }
                }
                for (int ii = 1; ii < kNB; ++ii) {
                    float v = x[ii];
                    for (int q = 0; q < ii; ++q) v -= pnl.at(ii, q) * x[q];
                    x[ii] = v;
                }
                for (int ii = 0; ii < kNB; ++ii) M.A(k + ii, c) = x[ii];
            }
        }
        // 4. first panel: the strip of the second panel's columns; second
        //    panel: the trailing matrix by both (K = 64)
        if (!second) {
            const int r1 = k + kNB, ms = n - r1;
            if (ms <= small_max) tile_update_small(M, pnl, r1, k, ms, std::min(kNB, ms), kNB);
            else tile_sweep(M, pnl, r1, k, ms, std::min(kNB, ms), kNB);
            continue;
        }
        if (!right) continue;
        const int mr = n - c1;
        if (mr <= small_max) tile_update_small(M, pnl, c1, k0, mr, mr, 2 * kNB);
        else tile_sweep(M, pnl, c1, k0, mr, mr, 2 * kNB);
    }
    // 5. the later blocks' interchanges applied to each earlier block's
    //    columns at the end: inv (rowmap) composed backwards; per 32
    //    columns, rows [kn, n) staged in the panel array, each moved row
    //    written to inv(row)
    const int blocks = (n + 2 * kNB - 1) / (2 * kNB);
    if (blocks > 1) {
        for (int r = 0; r < n; ++r) rowmap.at(0, r) = r;
        for (int b = blocks - 2; b >= 0; --b) {
            const int kn = (b + 1) * 2 * kNB;
            for (int j = std::min(kn + 2 * kNB, n) - 1; j >= kn; --j)
                std::swap(rowmap.at(0, j), rowmap.at(0, piv_all.at(0, j)));
            for (int h = 0; h < 2; ++h) {
                const int cb = b * 2 * kNB + h * kNB;
                for (int r = kn; r < n; ++r)
                    for (int c = 0; c < kNB; ++c) pnl.at(r - kn, c) = M.A(r, cb + c);
                for (int r = kn; r < n; ++r) {
                    const int d = rowmap.at(0, r);
                    if (d == r) continue;
                    for (int c = 0; c < kNB; ++c) M.A(d, cb + c) = pnl.at(r - kn, c);
                }
            }
        }
    }
}

template <typename T>
bool get(FILE* f, T* p, size_t count)
{
    return std::fread(p, sizeof(T), count, f) == count;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: %s input output [small_max]\n", argv[0]);
        return 2;
    }
    if (argc == 4) small_max = std::atoi(argv[3]);
    FILE* f = std::fopen(argv[1], "rb");
    if (f == nullptr) return 2;
    int batch = 0;
    long long total = 0;
    if (!get(f, &batch, 1) || !get(f, &total, 1) || batch <= 0 || total <= 0) return 2;
    std::vector<int> n(batch), lda(batch);
    std::vector<long long> off(batch), poff(batch);
    Checked M;
    M.flat.resize(total);
    if (!get(f, n.data(), batch) || !get(f, lda.data(), batch) || !get(f, off.data(), batch)
        || !get(f, poff.data(), batch) || !get(f, M.flat.data(), total))
        return 2;
    std::fclose(f);
    long long ptotal = 0;
    for (int v : n) ptotal += std::max(v, 1);
    M.ipiv.assign(ptotal, 0);

    // order_kernel: counting sort on n / 16, largest first
    std::vector<int> hist(kBuckets, 0), start(kBuckets, 0), order(batch, -1);
    auto bucket = [&](int i) {
        const int v = n[i] / 16;
        return kBuckets - 1 - (v < 0 ? 0 : (v >= kBuckets ? kBuckets - 1 : v));
    };
    for (int i = 0; i < batch; ++i) ++hist[bucket(i)];
    for (int b = 0, s = 0; b < kBuckets; ++b) {
        start[b] = s;
        s += hist[b];
    }
    for (int i = 0; i < batch; ++i) {
        const int pos = start[bucket(i)]++;
        if (pos < 0 || pos >= batch) M.report("order", pos, i);
        else order[pos] = i;
    }
    std::vector<int> info(batch, 99);
    for (int q = 0; q < batch; ++q) {
        const int id = order[q];
        if (id < 0 || id >= batch) {
            M.report("queue id", q, id);
            continue;
        }
        const int nn = n[id];
        if (nn <= 0 || nn > kLuMaxOrder || lda[id] < nn) {
            info[id] = nn == 0 ? 0 : -1;
            continue;
        }
        M.base = off[id];
        M.lda = lda[id];
        M.n = nn;
        M.lo = off[id];
        M.hi = off[id] + static_cast<long long>(lda[id]) * nn;
        M.plo = poff[id];
        M.phi = poff[id] + nn;
        info[id] = 0;
        lu_matrix(M, nn, &info[id]);
    }
    std::printf("batch %d: out-of-range accesses %lld\n", batch, M.violations);
    FILE* o = std::fopen(argv[2], "wb");
    if (o == nullptr) return 2;
    std::fwrite(M.flat.data(), sizeof(float), M.flat.size(), o);
    std::fwrite(M.ipiv.data(), sizeof(int), M.ipiv.size(), o);
    std::fwrite(info.data(), sizeof(int), info.size(), o);
    std::fclose(o);
    return M.violations != 0;
}
