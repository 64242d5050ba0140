// eri_part1.hip -- instantiates the ERI kernels of classes [4, 7).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part1, 4, 7)
