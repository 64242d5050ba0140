// eri_part5.hip -- instantiates the ERI kernels of classes [21, 30).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part5, 21, 30)
