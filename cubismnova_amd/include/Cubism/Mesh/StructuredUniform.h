// Mesh/StructuredUniform.h -- the drop-in's header under the reference's file name
// (include/Cubism/Mesh/StructuredUniform.h): application code that includes it gets the
// declarations of Cubism/AMD.h.
#pragma once
#include "Cubism/AMD.h"
