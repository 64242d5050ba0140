// eri_part2.hip -- instantiates the ERI kernels of classes [7, 10).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part2, 7, 10)
