// Host scene validation and packing (see gl_scene.cpp).
#pragma once

#include <string>

#include "../../include/gravlume.h"
#include "gl_types.h"

namespace gl {

// f64 view of a validated scene kept on the host.
struct SceneHost {
  gl_scene_desc desc;
  double r_plus;
  int extremality;  // 0 sub, 1 extremal, 2 super
  double g_tt;
  double u[4], right[4], up[4], arrival[4];
  double tan_half_fov;
  int up_axis;  // 0: the image-up hint; 1-3: the coordinate axis of the fallback
};

// Validates one scene and fills its device record; returns a gl_status and
// the reason in *err.
int prepare_scene(const gl_scene_desc& d, SceneDev* out, SceneHost* host, std::string* err);

// Exact extremality class of (M, a, Q): 0 sub-, 1 extremal, 2 super-extremal.
int extremality_f64(double mass, double spin, double charge);

}  // namespace gl
