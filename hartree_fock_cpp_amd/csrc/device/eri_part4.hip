// eri_part4.hip -- instantiates the ERI kernels of classes [47, 55).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part4, 47, 55)
