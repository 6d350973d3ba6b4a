// Split-FP16 kernels (limb split and FP16 MFMA contractions).
#include "device_common.hpp"
#include "kernels.hpp"

namespace bdc {}
