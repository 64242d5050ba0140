// eri_part6.hip -- instantiates the ERI kernels of classes [30, 42).
#include "eri_launch_impl.hpp"
HFX_DEFINE_PART(launch_eri_part6, 30, 42)
