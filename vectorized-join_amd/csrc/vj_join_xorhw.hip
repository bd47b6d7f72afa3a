// The join pipeline's hashing stages for custom_xor_hw.
#include "vj_join_stages.hpp"

template struct vj::StageLaunch<vj::HP_XOR_HW>;
