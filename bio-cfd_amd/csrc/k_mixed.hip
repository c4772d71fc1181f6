// mixed precision path (later)
#include "biocfd_rt.hpp"
