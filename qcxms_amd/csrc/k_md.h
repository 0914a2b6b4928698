// MD kernels (leap-frog, heating, connectivity, bookkeeping): see below.
#pragma once
#include "k_grad.h"
