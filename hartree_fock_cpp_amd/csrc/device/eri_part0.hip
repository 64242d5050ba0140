// eri_part0.hip -- instantiates the ERI kernels of classes [0, 4).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part0, 0, 4)
