// Core/Index.h -- the drop-in's header under the reference's file name
// (include/Cubism/Core/Index.h): application code that includes it gets the
// declarations of Cubism/AMD.h.
#pragma once
#include "Cubism/AMD.h"
