// eri_part4.hip -- instantiates the ERI kernels of classes [13, 21).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part4, 13, 21)
