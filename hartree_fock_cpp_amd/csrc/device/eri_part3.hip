// eri_part3.hip -- instantiates the ERI kernels of classes [38, 47).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part3, 38, 47)
