// The join pipeline's hashing stages for FNV-1a.
#include "vj_join_stages.hpp"

template struct vj::StageLaunch<vj::HP_FNV>;
