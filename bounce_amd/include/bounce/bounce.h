// bounce/bounce.h -- the drop-in API of the MI355X batched step (libbounce_amd.so + libbxstep.so).
//
// Programs written against the reference's <bounce/bounce.h> for rigid bodies (worlds, bodies,
// fixtures with sphere, capsule and hull shapes, box hulls, revolute joints) compile against this
// header unchanged; b3World::Step runs on the GPU.
#pragma once
#include "common/math.h"
#include "collision/shapes.h"
#include "dynamics/world.h"
