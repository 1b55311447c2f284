// kifi/kifi.hpp -- the C++ drop-in API of the MI355X fusion library: the
// reference's classes (KinectFission/kifi/*.h) over the C ABI of include/kifi.h.
// Link with libkifi_hip.so.
#pragma once

#include "kifi/DepthSensorParams.h"
#include "kifi/DepthStream.h"
#include "kifi/ICP.h"
#include "kifi/Integrator.h"
#include "kifi/Mesher.h"
#include "kifi/Pipeline.h"
#include "kifi/Renderer.h"
#include "kifi/Volume.h"
#include "kifi/Voxel.h"
#include "kifi/util/math.h"
#include "kifi/util/vector2d.h"
