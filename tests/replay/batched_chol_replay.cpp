// Host replay of the batched Cholesky kernel's address arithmetic
// (hygondoubleblockedfactorizations_amd/csrc/batched.hip: order_kernel,
// batched_kernel<kPrecFP32, false>, chol_matrix in 128-column outer blocks of
// four 32-column panels, tile_sweep<kPrecFP32, true> with its k slices and
// two operand buffers and, for trailing orders up to the small size class,
// tile_update_small<kPrecFP32, true>).
//
// Every phase between two workgroup barriers runs for tid = 0..255 (or lane
// = 0..63 of wave 0) in turn.  Every access to the packed matrix buffer and
// to the workgroup arrays goes through a checked accessor that counts
// accesses outside (a) the whole buffer, (b) the current matrix's lda x n
// block, (c) its n x n part, and writes above the diagonal.  Values are
// computed in FP32 in the kernel's order within each thread (the MFMA tiles'
// k order is the sequential one here), so info and the factor can be
// checked against LAPACK.
//
// Input (little endian): int32 batch, int64 total, int32 n[batch],
// int32 lda[batch], int64 offsets[batch], float flat[total].  Output: float
// flat[total], int32 info[batch].  Optional third argument: the small size
// class's largest trailing order (0 when absent).  Prints the count of
// out-of-range accesses; exit 1 if any.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr int kNB = 32, kT = 64, kMaxOrder = 4096, kBuckets = 256, kThreads = 256,
              kSmallT = 32, kPanels = 4;
constexpr int kTileFloats = 2 * kNB * (kT + 1);
constexpr int kSmallScratch = 4 * 2 * kNB * (kSmallT + 1);
constexpr int kScratch = kSmallScratch > 2 * kTileFloats ? kSmallScratch : 2 * kTileFloats;
int small_max = 0;

struct Checked {
    std::vector<float> flat;
    long long lo = 0, hi = 0;  // current matrix
    long long base = 0, lda = 0;
    int n = 0;
    long long violations = 0;
    float sink = 0.0f;

    void report(const char* what, long long a, long long b)
    {
        if (violations++ < 20) std::printf("out of range: %s (%lld, %lld)\n", what, a, b);
    }
    float& at(long long r, long long c)
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
    float get(long long r, long long c) { return at(r, c); }
    void put(long long r, long long c, float v)
    {
        if (c > r) report("write above the diagonal", r, c);
        at(r, c) = v;
    }
};

struct Shared {
    float lt[kNB][kNB];
    float rdiag[kNB];
    float scratch[kScratch];
    long long* violations;
    float& s(int i)
    {
        static float sink;
        if (i < 0 || i >= kScratch) {
            ++*violations;
            return sink;
        }
        return scratch[i];
    }
    float& l(int r, int c)
    {
        static float sink;
        if (r < 0 || r >= kNB || c < 0 || c >= kNB) {
            ++*violations;
            return sink;
        }
        return lt[r][c];
    }
};

int imin(int a, int b) { return a < b ? a : b; }

// C = A(r1.., r1..) m x nc (lower), operands P = A(r1:n, k:k+kw): tiles of
// 64 x 64 in the kernel's order, k slices of 32, operands staged in buffer
// (step % 2) of the scratch; each lane's C elements read at the first slice,
// written after the last.
void tile_sweep(Checked& M, Shared& S, int r1, int m, int nc, int k, int kw)
{
    const int tiles_r = (m + kT - 1) / kT, tiles_c = (nc + kT - 1) / kT;
    const int slices = (kw + kNB - 1) / kNB;
    std::vector<float> acc(static_cast<size_t>(kThreads) * 16), cv(acc.size());
    int step = 0;
    for (int tj = 0; tj < tiles_c; ++tj)
        for (int ti = tj; ti < tiles_r; ++ti)
            for (int sl = 0; sl < slices; ++sl, ++step) {
                const int r0 = ti * kT, c0 = tj * kT, k0 = sl * kNB;
                auto buf = [&](int half, int kk, int i) -> float& {
                    return S.s((step & 1) * kTileFloats + half * kNB * (kT + 1) + kk * (kT + 1) + i);
                };
                for (int tid = 0; tid < kThreads; ++tid)
                    for (int q = 0; q < kNB * kT / kThreads; ++q) {
                        const int e = tid + kThreads * q;
                        const int kk = k0 + e / kT, i = e % kT;
                        float va = 0.0f, vb = 0.0f;
                        if (kk < kw) {
                            if (r0 + i < m) va = M.get(r1 + r0 + i, k + kk);
                            if (c0 + i < nc) vb = M.get(r1 + c0 + i, k + kk);
                        }
                        buf(0, e / kT, i) = va;
                        buf(1, e / kT, i) = vb;
                    }
                for (int tid = 0; tid < kThreads; ++tid) {
                    const int lane = tid & 63, wave = tid >> 6, wm = wave & 1, wn = wave >> 1;
                    const int row = wm * 32 + (lane & 31);
                    const int gr = r0 + row;
                    for (int r = 0; r < 16; ++r) {
                        const int lc = wn * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                        const int gc = c0 + lc;
                        const bool ok = gc < nc && gr < m && gr >= gc;
                        float& a = acc[tid * 16 + r];
                        if (sl == 0) {
                            cv[tid * 16 + r] = ok ? M.get(r1 + gr, r1 + gc) : 0.0f;
                            a = 0.0f;
                        }
                        for (int kk = 0; kk < kNB; ++kk) a = std::fma(buf(1, kk, lc), buf(0, kk, row), a);
                        if (sl == slices - 1 && ok) M.put(r1 + gr, r1 + gc, cv[tid * 16 + r] - a);
                    }
                }
            }
}

// The small size class: wave w takes tiles w, w + 4, ... of the 32-blocks of
// the lower trapezoid (row-major), one 32-wide k slice at a time staged in
// its 2 x 32 x 33 slice of the scratch.
void tile_update_small(Checked& M, Shared& S, int r1, int m, int nc, int k, int kw)
{
    const int tr = (m + kSmallT - 1) / kSmallT, tc = (nc + kSmallT - 1) / kSmallT;
    const int full = tc < tr ? tc : tr;
    const int count = full * (full + 1) / 2 + (tr - full) * tc;
    for (int wave = 0; wave < 4; ++wave)
        for (int t = wave; t < count; t += 4) {
            int ti, tj;
            if (t < full * (full + 1) / 2) {
                ti = 0;
                while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
                tj = t - ti * (ti + 1) / 2;
            } else {
                const int u = t - full * (full + 1) / 2;
                ti = full + u / tc;
                tj = u % tc;
            }
            if (ti >= tr || tj >= tc || tj > ti) M.report("small tile index", ti, tj);
            const int r0 = ti * kSmallT, c0 = tj * kSmallT;
            std::vector<float> acc(64 * 16, 0.0f);
            auto slot = [&](int half, int kk, int i) -> float& {
                return S.s(wave * 2 * kNB * (kSmallT + 1) + half * kNB * (kSmallT + 1) + kk * (kSmallT + 1) + i);
            };
            for (int k0 = 0; k0 < kw; k0 += kNB) {
                for (int lane = 0; lane < 64; ++lane)
                    for (int e = lane; e < kNB * kSmallT; e += 64) {
                        const int i = e % kSmallT, kk = e / kSmallT;
                        float va = 0.0f, vb = 0.0f;
                        if (k0 + kk < kw) {
                            if (r0 + i < m) va = M.get(r1 + r0 + i, k + k0 + kk);
                            if (c0 + i < nc) vb = M.get(r1 + c0 + i, k + k0 + kk);
                        }
                        slot(0, kk, i) = va;
                        slot(1, kk, i) = vb;
                    }
                for (int lane = 0; lane < 64; ++lane) {
                    const int rc = lane & 31;
                    for (int r = 0; r < 16; ++r) {
                        const int lc = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                        for (int kk = 0; kk < kNB; ++kk)
                            acc[lane * 16 + r] = std::fma(slot(1, kk, lc), slot(0, kk, rc), acc[lane * 16 + r]);
                    }
                }
            }
            for (int lane = 0; lane < 64; ++lane) {
                const int gr = r0 + (lane & 31);
                if (gr >= m) continue;
                for (int r = 0; r < 16; ++r) {
                    const int gc = c0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    if (gc < nc && gr >= gc)
                        M.put(r1 + gr, r1 + gc, M.get(r1 + gr, r1 + gc) - acc[lane * 16 + r]);
                }
            }
        }
}

// one 32-column panel at column k (width kb): false on a failed pivot
bool panel(Checked& M, Shared& S, int n, int k, int kb, int* info)
{
    // POTF2 on wave 0: lane i holds row min(i, kb - 1), columns clamped to kb - 1
    std::vector<float> x(64 * kNB);
    for (int lane = 0; lane < 64; ++lane) {
        const int r = k + imin(lane, kb - 1);
        for (int c = 0; c < kNB; ++c) x[lane * kNB + c] = M.get(r, k + imin(c, kb - 1));
    }
    int bad = -1;
    for (int j = 0; j < kNB; ++j) {
        if (!(j < kb && bad < 0)) continue;
        const float djj = x[j * kNB + j];
        if (!(djj > 0.0f) || !std::isfinite(djj)) {
            bad = j;
            continue;
        }
        const float s = std::sqrt(djj);
        std::vector<float> col(64);
        for (int lane = 0; lane < 64; ++lane) {
            float& v = x[lane * kNB + j];
            v = lane == j ? s : (lane > j ? v / s : v);
        }
        for (int lane = 0; lane < 64; ++lane) col[lane] = x[lane * kNB + j];
        for (int lane = 0; lane < 64; ++lane)
            for (int c = j + 1; c < kNB; ++c) x[lane * kNB + c] -= col[lane] * col[c];
    }
    for (int lane = 0; lane < kNB; ++lane) {
        for (int c = 0; c < kNB; ++c) S.l(lane, c) = x[lane * kNB + c];
        S.rdiag[lane] = 1.0f / x[lane * kNB + lane];
    }
    if (bad >= 0) *info = k + bad + 1;
    for (int e = 0; e < kNB * kNB; ++e) {
        const int i = e % kNB, c = e / kNB;
        if (i < kb && c <= i) M.put(k + i, k + c, S.l(i, c));
    }
    if (bad >= 0) return false;
    // TRSM rows [k + kb, n), columns clamped to kb - 1 on the loads
    for (int r = k + kb; r < n; ++r) {
        float v[kNB];
        for (int j = 0; j < kNB; ++j) v[j] = M.get(r, k + imin(j, kb - 1));
        for (int j = 0; j < kNB; ++j) {
            float t = v[j];
            for (int c = 0; c < j; ++c) t -= v[c] * S.l(j, c);
            v[j] = t * S.rdiag[j];
        }
        for (int j = 0; j < kb; ++j) M.put(r, k + j, v[j]);
    }
    return true;
}

void update(Checked& M, Shared& S, int n, int r1, int nc, int k, int kw)
{
    const int m = n - r1;
    if (nc <= 0 || nc > m || kw <= 0 || kw > kPanels * kNB) M.report("update shape", nc, kw);
    if (m <= small_max) tile_update_small(M, S, r1, m, nc, k, kw);
    else tile_sweep(M, S, r1, m, nc, k, kw);
}

void chol_matrix(Checked& M, int n, int* info)
{
    static Shared S;
    S.violations = &M.violations;
    // outer blocks of kPanels panels: before panel q its columns by panels
    // 0 .. q - 1 of the block, after the last the trailing matrix
    for (int k = 0; k < n; k += kPanels * kNB) {
        const int kbb = imin(kPanels * kNB, n - k);
        const int nq = (kbb + kNB - 1) / kNB;
        for (int q = 0; q <= nq; ++q) {
            const int kp = imin(k + q * kNB, k + kbb);
            if (q > 0 && kp < n) {
                const bool strip = q < nq;
                update(M, S, n, kp, strip ? imin(kNB, k + kbb - kp) : n - kp, k, kp - k);
            }
            if (q < nq && !panel(M, S, n, kp, imin(kNB, k + kbb - kp), info)) return;
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
    std::vector<long long> off(batch);
    Checked M;
    M.flat.resize(total);
    if (!get(f, n.data(), batch) || !get(f, lda.data(), batch) || !get(f, off.data(), batch)
        || !get(f, M.flat.data(), total))
        return 2;
    std::fclose(f);

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
        if (nn <= 0 || nn > kMaxOrder || lda[id] < nn) {
            info[id] = nn == 0 ? 0 : -1;
            continue;
        }
        M.base = off[id];
        M.lda = lda[id];
        M.n = nn;
        M.lo = off[id];
        M.hi = off[id] + static_cast<long long>(lda[id]) * nn;
        info[id] = 0;
        chol_matrix(M, nn, &info[id]);
    }
    std::printf("batch %d: out-of-range accesses %lld\n", batch, M.violations);
    FILE* o = std::fopen(argv[2], "wb");
    if (o == nullptr) return 2;
    std::fwrite(M.flat.data(), sizeof(float), M.flat.size(), o);
    std::fwrite(info.data(), sizeof(int), info.size(), o);
    std::fclose(o);
    return M.violations != 0;
}
