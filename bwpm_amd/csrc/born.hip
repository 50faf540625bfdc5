// Born scattering at weak index fluctuations (device path added with the Born kernels).
#include "born.hpp"
