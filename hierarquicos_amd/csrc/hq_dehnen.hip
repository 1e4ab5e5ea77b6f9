// hq_dehnen.hip -- Dehnen dual-tree walk (octree_default.f90:1095-1712).
#include "hq_tree.h"

namespace hq {

int tree_forces_dehnen(hq_tree* t, double theta2, double G, double eps2) {
    (void)t; (void)theta2; (void)G; (void)eps2;
    return fail(HQ_ERR_ARGS, "Dehnen methods (20-22) are not available in this build");
}

}  // namespace hq
