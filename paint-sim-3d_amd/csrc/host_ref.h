// host_ref.h - host helpers of the reference's film module used by the
// step's host side (FP32 as in src/film/mod.rs).
#pragma once
#include "../../include/canvaz_film.h"

namespace czh {
float to_radians(float d);
float powi4(float x);
float height_at(const cz_deposit& d, float px, float py);
float sampled_volume(const cz_deposit& d, float lattice);
float sampled_volume_in(const cz_deposit& d, float lattice, float lo0, float lo1, float hi0, float hi1);
float inside_fraction(const cz_deposit& d, float w, float h);
float max_height(const cz_deposit& d);
}  // namespace czh
