// Launch-configuration tuner.
//
// The reference tunes only the thread block, online, one trial per loop
// execution, exploring by bisection of unseen intervals when DVMH_OPT_PARAMS=1
// (rts/dvmh/loop.cpp:1487-1909).  Here a candidate is a whole launch
// configuration (block, iterations per thread, XCD-aware order, kernel variant,
// LDS window / wavefront tile, layout transform).  Modes:
//   0  the default configuration (the reference's default blocks,
//      rts/dvmh/loop.cpp:1464-1485);
//   1  exhaustive: successive executions of a loop site try every candidate,
//      then the fastest is used; every timing is recorded;
//   2  cost model: the learned model ranks the candidates, the best few are
//      measured and the fastest kept (falls back to mode 1 without a model);
//   3  fixed: best recorded candidate from the tuning records.
// Configurations never change arithmetic, only the mapping of iterations.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <algorithm>

#include "rt.h"
#include "tuner.h"

namespace dvmh {

static std::map<std::string, SiteInfo> sites;

std::map<std::string, SiteInfo> &tunerSites() { return sites; }

static void addCand(std::vector<Candidate> &v, int bx, int by, int bz, std::initializer_list<std::pair<int, long>> t) {
    Candidate c;
    c.block[0] = bx; c.block[1] = by; c.block[2] = bz;
    for (int i = 0; i < DVMH_AMD_TUNE_N; i++) c.tune[i] = 0;
    c.tune[DVMH_AMD_T_DEPTH] = 1;
    for (auto &kv : t) c.tune[kv.first] = kv.second;
    for (auto &e : v)
        if (e.same(c)) return;
    v.push_back(c);
}

// Candidate lists.  Index 0 is always the default configuration.
std::vector<Candidate> makeCandidates(int kind, int loopRank, int depCount, const long extents[], long depMask) {
    std::vector<Candidate> v;
    int mapped = loopRank - (kind == DVMH_AMD_KIND_PARALLEL ? 0 : depCount);   // axes spread over threads
    const int tfMax = settings.tuneTransforms ? 1 : 0;   // also try device layout transforms
    if (kind == DVMH_AMD_KIND_PARALLEL || kind == DVMH_AMD_KIND_STENCIL || kind == DVMH_AMD_KIND_BATCH) {
        if (kind == DVMH_AMD_KIND_BATCH) {
            // batch kernels walk many small instances' tiles: wider rows and four
            // rows per thread (best of the 2-D sweep over 512 grids of sides
            // 64..1024; 1.3x over the per-loop default below)
            // (rank 2: a batch of stencil loops runs its handler's own march
            // shape instead, DVMH_AMD_T_OWN: 256 threads, 4 points each, 8 rows
            // per tile, 1.4x the direct tiles on the same sweep)
            if (loopRank >= 3) addCand(v, 64, 4, 1, {{DVMH_AMD_T_DEPTH, 4}});
            else if (loopRank == 2) addCand(v, 128, 4, 1, {{DVMH_AMD_T_DEPTH, 4}, {DVMH_AMD_T_OWN, 1}});
            else addCand(v, 256, 1, 1, {});
        }
        if (loopRank >= 3) addCand(v, 32, 4, 2, {});
        else if (loopRank == 2) addCand(v, 32, 8, 1, {});
        else addCand(v, 256, 1, 1, {});
        const int xs[] = {64, 128, 256};
        const int ys[] = {1, 2, 4, 8, 16};
        const int zs[] = {1, 2, 4};
        const int depths[] = {1, 2, 4, 8};
        for (int bx : xs)
            for (int by : ys)
                for (int bz : zs) {
                    if (loopRank < 3 && bz > 1) continue;
                    if (loopRank < 2 && by > 1) continue;
                    int n = bx * by * bz;
                    if (n < 64 || n > 1024) continue;
                    for (int dp : depths) {
                        if (loopRank < 2 && dp > 1) continue;
                        for (int xcd = 0; xcd <= 1; xcd++)
                            addCand(v, bx, by, bz, {{DVMH_AMD_T_DEPTH, dp}, {DVMH_AMD_T_XCD, xcd}});
                    }
                }
        if ((kind == DVMH_AMD_KIND_STENCIL || kind == DVMH_AMD_KIND_BATCH) && (loopRank == 2 || loopRank == 3)) {
            // (a batch of stencil loops runs the same march per tile; loops without
            // a stencil array ignore the variant and run the direct tiles)
            // LDS march variant: block (bx, by), TILE_A points per thread on the
            // innermost axis, CHUNK planes per workgroup; three planes of
            // (by + 2) x (bx * TILE_A + 2) FP64 elements must fit 64 KB of LDS
            const int mxs[] = {64, 128, 256, 512, 1024};
            const int mys[] = {1, 2, 4, 8};
            const int mcs[] = {1, 2, 3, 4};
            const int marches[] = {8, 16, 32, 64};
            for (int bx : mxs)
                for (int by : mys) {
                    if (loopRank == 2 && by > 1) continue;
                    if (loopRank == 3 && bx > 256) continue;
                    int n = bx * by;
                    if (n < 64 || n > 1024) continue;
                    for (int c : mcs) {
                        long rows = loopRank == 3 ? by + 2 : 1;
                        if (3 * rows * ((long)bx * c + 2) * 8 > 64512) continue;
                        for (int m : marches)
                            for (int xcd = 0; xcd <= 1; xcd++)
                                addCand(v, bx, by, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, c}, {DVMH_AMD_T_CHUNK, m},
                                                       {DVMH_AMD_T_XCD, xcd}});
                    }
                }
        }
    } else if (kind == DVMH_AMD_KIND_LINE) {
        // threads over the independent axes, each walks its line
        if (mapped >= 2) addCand(v, 32, 8, 1, {});
        else addCand(v, 256, 1, 1, {});
        const int xs[] = {32, 64, 128, 256};
        const int ys[] = {1, 2, 4, 8};
        for (int bx : xs)
            for (int by : ys) {
                if (mapped < 2 && by > 1) continue;
                int n = bx * by;
                if (n < 64 || n > 1024) continue;
                for (int xcd = 0; xcd <= 1; xcd++)
                    for (int tf = 0; tf <= tfMax; tf++)
                        addCand(v, bx, by, 1, {{DVMH_AMD_T_XCD, xcd}, {DVMH_AMD_T_TRANSFORM, tf}});
            }
        if (depMask > 1 && __builtin_popcountl((unsigned long)depMask) == 1 && (loopRank == 2 || loopRank == 3)) {
            // register-staged line sweep (dependent axis not innermost): threads
            // over the independent axes, TILE_A steps per chunk held in registers
            const int rbs[][2] = {{64, 1}, {128, 1}, {256, 1}, {64, 2}, {32, 4}};
            const int rcks[] = {8, 16};
            for (auto &bs : rbs) {
                if (mapped < 2 && bs[1] > 1) continue;
                for (int ck : rcks)
                    for (int xcd = 0; xcd <= 1; xcd++)
                        addCand(v, bs[0], bs[1], 1, {{DVMH_AMD_T_VARIANT, 2}, {DVMH_AMD_T_TILE_A, ck}, {DVMH_AMD_T_XCD, xcd}});
            }
        }
        if (depMask == 1 && (loopRank == 2 || loopRank == 3)) {
            // tiled line sweep (dependent axis innermost): blockDim.x lines per
            // workgroup, TILE_A iterations per staged chunk; LDS of two chunks
            // of lines x ((chunk + 2) | 1) FP64 slots plus a row offset per line
            const int tls[] = {64, 128, 256};
            const int cks[] = {8, 16, 32};
            for (int tl : tls)
                for (int ck : cks) {
                    if ((long)tl * (2 * ((ck + 2) | 1) * 8 + 8) + 16 > 64512) continue;
                    for (int reg = 0; reg <= 1; reg++)   // TILE_B 1: the walk runs on a register copy of the line's chunk
                        addCand(v, tl, 1, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, ck}, {DVMH_AMD_T_TILE_B, reg}});
                }
            // streamed line sweep (variant 3): a wave's 64 lines move in aligned
            // 128-byte chunks, coalesced, transposed through the wave's LDS rows
            for (int tl : tls)
                addCand(v, tl, 1, 1, {{DVMH_AMD_T_VARIANT, 3}});
        }
    } else if (kind == DVMH_AMD_KIND_PAIR) {
        // a fused pair runs the LDS march of its stencil loop, the second loop
        // inside it off the tile edges and on the edges in a second kernel:
        // taller tiles leave fewer edge points.  Default: rank 2 -- the
        // handler's own shape (DVMH_AMD_T_OWN): the widest of 256 x 4 points x
        // 32 planes (best of the sweep over 512 grids of sides 64..1024, 1.19
        // ms per time step) down to 64 x 1 x 8 that still gives every CU four
        // tiles; rank 3 -- 64 x 4 threads x 2 points, 32 planes.
        if (loopRank == 3) addCand(v, 64, 4, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, 2}, {DVMH_AMD_T_CHUNK, 32}});
        else addCand(v, 256, 1, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, 4}, {DVMH_AMD_T_CHUNK, 32}, {DVMH_AMD_T_OWN, 1}});
        const int mxs[] = {64, 128, 256, 512, 1024};
        const int mys[] = {1, 2, 4, 8};
        const int mcs[] = {1, 2, 4};
        const int marches[] = {8, 16, 32, 64};
        for (int bx : mxs)
            for (int by : mys) {
                if (loopRank == 2 && by > 1) continue;
                if (loopRank == 3 && (by < 2 || bx > 256)) continue;   // an edge row per tile row otherwise
                int n = bx * by;
                if (n < 64 || n > 1024) continue;
                for (int c : mcs) {
                    long rows = loopRank == 3 ? by + 2 : 1;
                    if (3 * rows * ((long)bx * c + 2) * 8 > 64512) continue;
                    for (int m : marches)
                        for (int xcd = 0; xcd <= 1; xcd++)
                            addCand(v, bx, by, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, c}, {DVMH_AMD_T_CHUNK, m},
                                                   {DVMH_AMD_T_XCD, xcd}});
                }
            }
    } else if (kind == DVMH_AMD_KIND_CHAIN) {
        // a chain of sweeps of a tiled-wavefront loop: tile shapes only (the
        // chain handler runs nothing else); times are per sweep of the chain
        addCand(v, 256, 1, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, 8}, {DVMH_AMD_T_TILE_B, 8}});
        const int nts[] = {128, 256, 512};
        const int tas[] = {8, 12, 16};
        const int tcs[] = {8, 16, 32};
        for (int nt : nts)
            for (int ta : tas)
                for (int tc : tcs) {
                    if ((long)(ta + 2) * (ta + 2) * (tc + 2) * 8 > 64512) continue;
                    if (nt > 4 * ta * ta) continue;
                    addCand(v, nt, 1, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, ta}, {DVMH_AMD_T_TILE_B, tc}});
                }
    } else {
        // one launch per hyperplane: threads over the free dependent axes (y only
        // with three or more dependent axes), optional diagonal layout
        addCand(v, 64, 1, 1, {});
        const int xs[] = {32, 64, 128, 256};
        const int ys[] = {1, 2, 4, 8};
        for (int bx : xs)
            for (int by : ys) {
                if (depCount < 3 && by > 1) continue;
                int n = bx * by;
                if (n < 64 || n > 512) continue;
                for (int tf = 0; tf <= tfMax; tf++) addCand(v, bx, by, 1, {{DVMH_AMD_T_TRANSFORM, tf}});
            }
        if (loopRank == 3 && depCount == 3) {
            // tiled variant: TILE_A x TILE_A x TILE_B tiles staged in LDS with a
            // halo of one, one launch per tile-plane
            const int nts[] = {64, 128, 256, 512};
            const int tas[] = {4, 8, 12, 16};
            const int tcs[] = {8, 16, 32, 64};
            for (int nt : nts)
                for (int ta : tas)
                    for (int tc : tcs) {
                        if ((long)(ta + 2) * (ta + 2) * (tc + 2) * 8 > 64512) continue;
                        if (nt > 4 * ta * ta) continue;   // threads cover a tile's (a, b) columns
                        addCand(v, nt, 1, 1, {{DVMH_AMD_T_VARIANT, 1}, {DVMH_AMD_T_TILE_A, ta}, {DVMH_AMD_T_TILE_B, tc}});
                    }
        }
    }
    (void)extents;
    return v;
}

std::string siteKey(const Loop *l, int kind) {
    std::ostringstream k;
    k << l->file << ":" << l->line << "|k" << kind << "|r" << l->rank << "|m" << l->depMask << "|e";
    for (int i = 0; i < l->rank; i++) {
        long c = l->step[i] > 0 ? (l->hi[i] - l->lo[i]) / l->step[i] + 1 : (l->lo[i] - l->hi[i]) / (-l->step[i]) + 1;
        k << (i ? "x" : "") << c;
    }
    return k.str();
}

static SiteInfo &siteFor(SpecLoop *sl, int kind) {
    std::string key = siteKey(sl->loop, kind);
    auto it = sites.find(key);
    if (it != sites.end()) return it->second;
    SiteInfo &s = sites[key];
    s.key = key;
    s.kind = kind;
    s.rank = sl->loop->rank;
    s.depCount = __builtin_popcountl((unsigned long)sl->loop->depMask);
    s.depMask = sl->loop->depMask;
    for (int i = 0; i < sl->loop->rank && i < 8; i++) {
        Loop *l = sl->loop;
        s.extents[i] = l->step[i] > 0 ? (l->hi[i] - l->lo[i]) / l->step[i] + 1 : (l->lo[i] - l->hi[i]) / (-l->step[i]) + 1;
    }
    s.cands = makeCandidates(kind, s.rank, s.depCount, s.extents, s.depMask);
    s.times.assign(s.cands.size(), std::vector<double>());
    return s;
}

double SiteInfo::measured(int c) const {
    const auto &t = times[c];
    if (t.empty()) return -1;
    // the minimum of the repeats: the configuration's cost without interference
    return *std::min_element(t.begin(), t.end());
}

int SiteInfo::bestMeasured() const {
    int best = -1;
    double bt = 0;
    for (size_t c = 0; c < cands.size(); c++) {
        double t = measured((int)c);
        if (t < 0) continue;
        if (best < 0 || t < bt) { best = (int)c; bt = t; }
    }
    return best;
}

extern int modelRank(SiteInfo &s, std::vector<int> &order);   // cost model (model.cpp)

// Configurations forced per loop kind (tests, experiments): used instead of the
// mode's choice at every site of that kind; never timed or recorded.
static std::map<long, Candidate> forced;

// Candidates the cost-model mode times: the default and the model's best K.
// The default, the model's best K, and the model's best candidate of every
// kernel variant the site offers that the top K miss: a variant the records
// hold little of (so the model ranks it poorly) is still timed once.
std::vector<int> SiteInfo::trialSet() const {
    const int K = 4;
    std::vector<int> trial{0};
    for (int i = 0; i < (int)modelOrder.size() && (int)trial.size() < K + 1; i++)
        if (modelOrder[i] != 0) trial.push_back(modelOrder[i]);
    for (int c : modelOrder) {
        const long v = cands[c].tune[DVMH_AMD_T_VARIANT];
        bool seen = false;
        for (int t : trial) seen = seen || cands[t].tune[DVMH_AMD_T_VARIANT] == v;
        if (!seen) trial.push_back(c);
    }
    return trial;
}

// Candidates of the site still lacking their timed repeats in the current mode.
long SiteInfo::pending() const {
    long p = 0;
    if (settings.tuneMode == 2 && modelRanked) {
        for (int c : trialSet())
            if ((int)times[c].size() < settings.tuneRepeats) p++;
        return p;
    }
    if (settings.tuneMode != 1 && settings.tuneMode != 2) return 0;
    for (size_t c = 0; c < cands.size(); c++)
        if ((int)times[c].size() < settings.tuneRepeats) p++;
    return p;
}

void tunerPick(SpecLoop *sl, long kind, TuneChoice &out) {
    Loop *l = sl->loop;
    int chosen = 0;
    SiteInfo *s = nullptr;
    if (l->userBlockFlag) {
        // a user cuda_block wins over everything (rts/dvmh/loop.cpp:2141-2147)
        std::vector<Candidate> c = makeCandidates((int)kind, l->rank, __builtin_popcountl((unsigned long)l->depMask), nullptr,
                l->depMask);
        Candidate d = c[0];
        for (int i = 0; i < 3; i++) out.block[i] = l->userBlock[i];
        for (int i = 0; i < DVMH_AMD_TUNE_N; i++) out.tune[i] = d.tune[i];
        out.candidate = -1;
        return;
    }
    auto fit = forced.find(kind);
    if (fit != forced.end()) {
        for (int i = 0; i < 3; i++) out.block[i] = fit->second.block[i];
        for (int i = 0; i < DVMH_AMD_TUNE_N; i++) out.tune[i] = fit->second.tune[i];
        out.candidate = -1;
        return;
    }
    s = &siteFor(sl, (int)kind);
    int mode = settings.tuneMode;
    if (mode == 2 && !s->modelRanked) {
        s->modelOrder.clear();
        if (modelRank(*s, s->modelOrder) > 0) s->modelRanked = true;
        else mode = 1;
    }
    if (mode == 0) {
        chosen = 0;
    } else if (mode == 3) {
        int b = s->bestMeasured();
        chosen = b >= 0 ? b : 0;
    } else if (mode == 1 || (mode == 2 && !s->modelRanked)) {
        // next candidate lacking its repeats, else the best
        chosen = -1;
        for (size_t c = 0; c < s->cands.size(); c++)
            if ((int)s->times[c].size() < settings.tuneRepeats) { chosen = (int)c; break; }
        if (chosen < 0) chosen = s->bestMeasured();
    } else {
        // model: measure the top-K predictions (plus the default), then keep the best
        chosen = -1;
        std::vector<int> trial = s->trialSet();
        for (int c : trial)
            if ((int)s->times[c].size() < settings.tuneRepeats) { chosen = c; break; }
        if (chosen < 0) {
            double bt = 0;
            for (int c : trial) {
                double t = s->measured(c);
                if (t >= 0 && (chosen < 0 || t < bt)) { chosen = c; bt = t; }
            }
        }
    }
    // a candidate still lacking its timed repeats is being explored: time this
    // execution from here (before the handler launches anything) to its end
    if ((settings.tuneMode == 1 || settings.tuneMode == 2) && (int)s->times[chosen].size() < settings.tuneRepeats &&
            !graphCapturing() && sl->stream) {
        sl->explore = true;
        sl->tuneStart = tunerEvent(0);
        HIP_CHECK(hipEventRecord(sl->tuneStart, sl->stream));
    }
    const Candidate &c = s->cands[chosen];
    for (int i = 0; i < 3; i++) out.block[i] = c.block[i];
    for (int i = 0; i < DVMH_AMD_TUNE_N; i++) out.tune[i] = c.tune[i];
    out.candidate = chosen;
    s->lastPicked = chosen;
    sl->candidate = chosen;
    sl->siteKey = s->key;
}

// Two reusable events for timing explored executions (one execution at a time).
hipEvent_t tunerEvent(int which) {
    static hipEvent_t ev[2] = {nullptr, nullptr};
    if (!ev[which]) HIP_CHECK(hipEventCreate(&ev[which]));
    return ev[which];
}

void tunerRecord(SpecLoop *sl, double seconds) {
    auto it = sites.find(sl->siteKey);
    if (it == sites.end() || sl->candidate < 0) return;
    SiteInfo &s = it->second;
    if (sl->candidate < (int)s.times.size()) s.times[sl->candidate].push_back(seconds);
}

// ---------------------------------------------------------------- records
// One line per measurement: key \t candidate \t bx by bz \t tune[8] \t seconds
long tunerSave(const std::string &path) {
    std::ofstream f(path);
    if (!f) return -1;
    long n = 0;
    for (auto &kv : sites) {
        const SiteInfo &s = kv.second;
        for (size_t c = 0; c < s.cands.size(); c++)
            for (double t : s.times[c]) {
                f << s.key << "\t" << c << "\t" << s.cands[c].block[0] << " " << s.cands[c].block[1] << " " << s.cands[c].block[2] << "\t";
                for (int i = 0; i < DVMH_AMD_TUNE_N; i++) f << (i ? " " : "") << s.cands[c].tune[i];
                f << "\t" << t << "\n";
                n++;
            }
    }
    return n;
}

// Parse a key back into the site description (kind, rank, depCount, extents).
static bool parseKey(const std::string &key, SiteInfo &s) {
    size_t k = key.find("|k");
    if (k == std::string::npos) return false;
    s.kind = atoi(key.c_str() + k + 2);
    size_t r = key.find("|r", k);
    size_t m = key.find("|m", k);
    size_t e = key.find("|e", k);
    if (r == std::string::npos || m == std::string::npos || e == std::string::npos) return false;
    s.rank = atoi(key.c_str() + r + 2);
    long mask = atol(key.c_str() + m + 2);
    s.depCount = __builtin_popcountl((unsigned long)mask);
    s.depMask = mask;
    std::string ex = key.substr(e + 2);
    std::stringstream ss(ex);
    std::string part;
    int i = 0;
    while (std::getline(ss, part, 'x') && i < 8) s.extents[i++] = atol(part.c_str());
    return true;
}

long tunerLoad(const std::string &path) {
    std::ifstream f(path);
    if (!f) return -1;
    std::string line;
    long n = 0;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        std::string key, cs, bs, ts, secs;
        if (!std::getline(ss, key, '\t') || !std::getline(ss, cs, '\t') || !std::getline(ss, bs, '\t') || !std::getline(ss, ts, '\t') ||
                !std::getline(ss, secs, '\t'))
            continue;
        auto it = sites.find(key);
        if (it == sites.end()) {
            SiteInfo s;
            s.key = key;
            if (!parseKey(key, s)) continue;
            s.cands = makeCandidates(s.kind, s.rank, s.depCount, s.extents, s.depMask);
            s.times.assign(s.cands.size(), std::vector<double>());
            it = sites.emplace(key, s).first;
        }
        // match the candidate by its configuration (the list depends on settings)
        Candidate want;
        std::stringstream bss(bs), tss(ts);
        for (int i = 0; i < 3; i++) bss >> want.block[i];
        for (int i = 0; i < DVMH_AMD_TUNE_N; i++) tss >> want.tune[i];
        if (!bss || !tss) continue;
        const auto &cands = it->second.cands;
        for (size_t c = 0; c < cands.size(); c++)
            if (cands[c].same(want)) {
                it->second.times[c].push_back(atof(secs.c_str()));
                n++;
                break;
            }
    }
    return n;
}

void modelInit();   // model.cpp

void tunerInit() {
    if (!settings.tuneDb.empty()) tunerLoad(settings.tuneDb);
    modelInit();
}

void tunerShutdown() {
    if (!settings.tuneDb.empty() && settings.tuneMode != 0 && settings.tuneMode != 3) tunerSave(settings.tuneDb);
}

}  // namespace dvmh

using namespace dvmh;

extern "C" void dvmh_amd_tuner_set_mode_C(DvmType mode) {
    runtimeInit();
    DVMH_CHECK(mode >= 0 && mode <= 3, "Tuner mode must be 0..3");
    settings.tuneMode = (int)mode;
}

extern "C" void dvmh_amd_tuner_set_repeats_C(DvmType repeats) {
    runtimeInit();
    DVMH_CHECK(repeats >= 1, "Tuner repeats must be >= 1");
    settings.tuneRepeats = (int)repeats;
}

extern "C" void dvmh_amd_tuner_force_C(DvmType kind, const DvmType block[3], const DvmType tune[DVMH_AMD_TUNE_N]) {
    runtimeInit();
    if (!block) {
        forced.erase(kind);
        return;
    }
    Candidate c;
    for (int i = 0; i < 3; i++) c.block[i] = (int)block[i];
    for (int i = 0; i < DVMH_AMD_TUNE_N; i++) c.tune[i] = tune ? tune[i] : 0;
    forced[kind] = c;
}

extern "C" DvmType dvmh_amd_tuner_save_C(const char *path) { runtimeInit(); return tunerSave(path); }
extern "C" DvmType dvmh_amd_tuner_load_C(const char *path) { runtimeInit(); return tunerLoad(path); }
extern "C" void dvmh_amd_tuner_clear_C(void) { runtimeInit(); sites.clear(); }

extern "C" DvmType dvmh_amd_tuner_pending_C(void) {
    runtimeInit();
    long pending = 0;
    for (auto &kv : sites) pending += kv.second.pending();
    return pending;
}

namespace dvmh {
// Candidates still unmeasured at the sites of one source loop (file, line).
long tunerSitePending(const std::string &file, long line) {
    std::string prefix = file + ":" + std::to_string(line) + "|";
    long pending = 0;
    for (auto &kv : sites) {
        if (kv.first.compare(0, prefix.size(), prefix) != 0) continue;
        pending += kv.second.pending();
    }
    return pending;
}

// The same without the sites of one kind (a chain's own site, say).
long tunerSitePendingExcept(const std::string &file, long line, int kind) {
    std::string prefix = file + ":" + std::to_string(line) + "|";
    long pending = 0;
    for (auto &kv : sites) {
        if (kv.first.compare(0, prefix.size(), prefix) != 0 || kv.second.kind == kind) continue;
        pending += kv.second.pending();
    }
    return pending;
}
}  // namespace dvmh

extern "C" DvmType dvmh_amd_tuner_site_pending_C(const char *file, DvmType line) { runtimeInit(); return tunerSitePending(file, line); }

static DvmType tunerQuery(const char *file, DvmType line, int kind, DvmType out[13]) {
    std::string prefix = std::string(file) + ":" + std::to_string(line) + "|";
    for (auto &kv : sites) {
        if (kv.first.compare(0, prefix.size(), prefix) != 0) continue;
        // a loop's own sites, or the chain or pair site asked for
        const bool own = kv.second.kind != DVMH_AMD_KIND_CHAIN && kv.second.kind != DVMH_AMD_KIND_PAIR;
        if (kind < 0 ? !own : kv.second.kind != kind) continue;
        int b = kv.second.bestMeasured();
        if (b < 0) return 0;
        const Candidate &c = kv.second.cands[b];
        out[0] = b;
        out[1] = (DvmType)(kv.second.measured(b) * 1e9);
        out[2] = c.block[0]; out[3] = c.block[1]; out[4] = c.block[2];
        for (int i = 0; i < DVMH_AMD_TUNE_N; i++) out[5 + i] = c.tune[i];
        return 1;
    }
    return 0;
}

extern "C" DvmType dvmh_amd_tuner_query_C(const char *file, DvmType line, DvmType out[13]) {
    runtimeInit();
    return tunerQuery(file, line, -1, out);
}

// How the layout-transform candidates of a loop's sites fared (summed over its
// sites other than a chain's): out[0] transform candidates, out[1] of them
// timed, out[2] best timed transform candidate in ns (-1 if none), out[3] best
// timed natural-layout candidate in ns (-1 if none).  Returns the sites seen.
extern "C" DvmType dvmh_amd_tuner_transform_stats_C(const char *file, DvmType line, DvmType out[4]) {
    runtimeInit();
    std::string prefix = std::string(file) + ":" + std::to_string(line) + "|";
    out[0] = out[1] = 0;
    out[2] = out[3] = -1;
    long seen = 0;
    for (auto &kv : sites) {
        if (kv.first.compare(0, prefix.size(), prefix) != 0 || kv.second.kind == DVMH_AMD_KIND_CHAIN ||
                kv.second.kind == DVMH_AMD_KIND_PAIR) continue;
        seen++;
        const SiteInfo &si = kv.second;
        for (size_t c = 0; c < si.cands.size(); c++) {
            const bool tf = si.cands[c].tune[DVMH_AMD_T_TRANSFORM] != 0;
            if (tf) out[0]++;
            const double t = si.measured((int)c);
            if (t < 0) continue;
            if (tf) out[1]++;
            DvmType &best = out[tf ? 2 : 3];
            const DvmType ns = (DvmType)(t * 1e9);
            if (best < 0 || ns < best) best = ns;
        }
    }
    return seen;
}

extern "C" DvmType dvmh_amd_tuner_query_chain_C(const char *file, DvmType line, DvmType out[13]) {
    runtimeInit();
    return tunerQuery(file, line, DVMH_AMD_KIND_CHAIN, out);
}

extern "C" DvmType dvmh_amd_tuner_query_pair_C(const char *file, DvmType line, DvmType out[13]) {
    runtimeInit();
    return tunerQuery(file, line, DVMH_AMD_KIND_PAIR, out);
}
