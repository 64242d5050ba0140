// eri_part7.hip -- instantiates the ERI kernels of classes [42, 55).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part7, 42, 55)
