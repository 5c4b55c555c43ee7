// batch.h — host-side include of the batched render's device records (dev_batch.h): many
// (scene, camera, size) jobs served by one persistent launch that pulls tiles from a global
// atomic queue.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_batch.h"
