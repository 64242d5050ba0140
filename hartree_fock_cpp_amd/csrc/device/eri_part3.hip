// eri_part3.hip -- instantiates the ERI kernels of classes [10, 13).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part3, 10, 13)
