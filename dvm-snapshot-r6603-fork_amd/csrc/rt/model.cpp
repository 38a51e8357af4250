// Learned cost model hook of the tuner.
#include "tuner.h"

namespace dvmh {

int modelRank(SiteInfo &s, std::vector<int> &order) {
    (void)s;
    order.clear();
    return 0;
}

}  // namespace dvmh
