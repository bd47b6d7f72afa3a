// The join pipeline's hashing stages for every other hash kind (the generic word-stream path).
#include "vj_join_stages.hpp"

template struct vj::StageLaunch<vj::HP_GENERIC>;
