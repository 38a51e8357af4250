// Tuner data shared by the tuner (tuner.cpp) and the learned cost model (model.cpp).
#pragma once

#include <string>
#include <vector>

#include "rt.h"

namespace dvmh {

struct Candidate {
    int block[3];
    DvmType tune[DVMH_AMD_TUNE_N];
    bool same(const Candidate &o) const {
        for (int i = 0; i < 3; i++) if (block[i] != o.block[i]) return false;
        for (int i = 0; i < DVMH_AMD_TUNE_N; i++) if (tune[i] != o.tune[i]) return false;
        return true;
    }
};

// One loop site (source position + shape) on the GPU.
struct SiteInfo {
    std::string key;
    int kind = 0, rank = 0, depCount = 0;
    long depMask = 0;
    long extents[8] = {0};
    std::vector<Candidate> cands;
    std::vector<std::vector<double>> times;   // seconds per measured execution
    std::vector<int> modelOrder;              // candidates ranked by the cost model
    bool modelRanked = false;
    int lastPicked = -1;
    double measured(int c) const;
    int bestMeasured() const;
    std::vector<int> trialSet() const;   // cost-model mode: default + the model's best few
    long pending() const;                // candidates still to time in the current mode
};

// depMask: the loop's dependency mask (bit rank-1-a for axis a), -1 if unknown
std::vector<Candidate> makeCandidates(int kind, int loopRank, int depCount, const long extents[], long depMask = -1);
std::map<std::string, SiteInfo> &tunerSites();
long tunerSave(const std::string &path);
long tunerLoad(const std::string &path);

}  // namespace dvmh
