// optk_signatures.h -- the surface-kind signatures of the systems the trace
// kernels are specialised for (not part of the C ABI).
//
// A signature lists, for every surface of a system (the object surface
// first, the sensor last), the kinds its record selects: transformation,
// sag, sag transformation, material, mirror flag, rulings, line spacing,
// aperture, aperture transformation and aperture flags (kinds_pack).  A
// system whose records carry exactly these kinds and features runs a kernel
// with the kinds fixed at compile time; the parameters (radii, conics,
// matrices, tables) stay run-time data, so every perturbed or rescaled
// instance of a layout shares its kernel.  Any other system runs the
// generic kernels, which read the kinds from the records.
#pragma once

#include "optk_surface.h"

namespace optk {

#define OPTK_K_PLAIN_OBJECT                                                                     \
  kinds_pack(OPTK_XF_IDENTITY, OPTK_SAG_NONE, OPTK_XF_IDENTITY, OPTK_MATERIAL_VACUUM, 0,         \
             OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_NONE, OPTK_XF_IDENTITY, 0)

// the Newtonian telescope: object plane, conic primary mirror (translated,
// circular aperture), flat fold mirror and sensor (rotated and translated,
// rectangular apertures)
#define OPTK_SIGNATURE_NEWTONIAN                                                                 \
  OPTK_K_PLAIN_OBJECT,                                                                          \
      kinds_pack(OPTK_XF_TRANSLATE, OPTK_SAG_CONIC, OPTK_XF_IDENTITY, OPTK_MATERIAL_MIRROR, 1,   \
                 OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_CIRCLE, OPTK_XF_IDENTITY, \
                 OPTK_AP_ACTIVE),                                                               \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_NONE, OPTK_XF_IDENTITY, OPTK_MATERIAL_MIRROR, 1,       \
                 OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_RECTANGLE,              \
                 OPTK_XF_IDENTITY, OPTK_AP_ACTIVE),                                             \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_NONE, OPTK_XF_IDENTITY, OPTK_MATERIAL_VACUUM, 0,       \
                 OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_RECTANGLE,              \
                 OPTK_XF_IDENTITY, OPTK_AP_ACTIVE)

// the EUV spectrograph: object plane, off-axis conic primary with a
// multilayer (decentred sag, circular aperture), toroidal grating with a
// multilayer and polynomial line spacing (rectangular aperture), polygonal
// field stop, sensor
#define OPTK_SIGNATURE_SPECTROGRAPH                                                              \
  OPTK_K_PLAIN_OBJECT,                                                                          \
      kinds_pack(OPTK_XF_TRANSLATE, OPTK_SAG_CONIC, OPTK_XF_TRANSLATE,                           \
                 OPTK_MATERIAL_MULTILAYER_MIRROR, 1, OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT,   \
                 OPTK_APERTURE_CIRCLE, OPTK_XF_IDENTITY, OPTK_AP_ACTIVE),                       \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_TOROID, OPTK_XF_IDENTITY,                              \
                 OPTK_MATERIAL_MULTILAYER_MIRROR, 1, OPTK_RULINGS_IDEAL, OPTK_SPACING_POLY1D,    \
                 OPTK_APERTURE_RECTANGLE, OPTK_XF_IDENTITY, OPTK_AP_ACTIVE),                    \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_NONE, OPTK_XF_IDENTITY, OPTK_MATERIAL_VACUUM, 0,       \
                 OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_POLYGON,                \
                 OPTK_XF_IDENTITY, OPTK_AP_ACTIVE),                                             \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_NONE, OPTK_XF_IDENTITY, OPTK_MATERIAL_VACUUM, 0,       \
                 OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_RECTANGLE,              \
                 OPTK_XF_IDENTITY, OPTK_AP_ACTIVE)

// the EUV spectrograph with a tilted primary (its transformation a full
// affine map): every instance of the tolerance batch (decentred and tilted
// surfaces) has these kinds
#define OPTK_SIGNATURE_SPECTROGRAPH_TILTED                                                       \
  OPTK_K_PLAIN_OBJECT,                                                                          \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_CONIC, OPTK_XF_TRANSLATE,                              \
                 OPTK_MATERIAL_MULTILAYER_MIRROR, 1, OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT,   \
                 OPTK_APERTURE_CIRCLE, OPTK_XF_IDENTITY, OPTK_AP_ACTIVE),                       \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_TOROID, OPTK_XF_IDENTITY,                              \
                 OPTK_MATERIAL_MULTILAYER_MIRROR, 1, OPTK_RULINGS_IDEAL, OPTK_SPACING_POLY1D,    \
                 OPTK_APERTURE_RECTANGLE, OPTK_XF_IDENTITY, OPTK_AP_ACTIVE),                    \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_NONE, OPTK_XF_IDENTITY, OPTK_MATERIAL_VACUUM, 0,       \
                 OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_POLYGON,                \
                 OPTK_XF_IDENTITY, OPTK_AP_ACTIVE),                                             \
      kinds_pack(OPTK_XF_AFFINE, OPTK_SAG_NONE, OPTK_XF_IDENTITY, OPTK_MATERIAL_VACUUM, 0,       \
                 OPTK_RULINGS_NONE, OPTK_SPACING_CONSTANT, OPTK_APERTURE_RECTANGLE,              \
                 OPTK_XF_IDENTITY, OPTK_AP_ACTIVE)

constexpr int SIGNATURE_MAX_SURFACES = 8;

struct Signature {
  int features;  // the kernel variant (FEAT_*) the system needs
  int num_surfaces;
  uint32_t kinds[SIGNATURE_MAX_SURFACES];
};

// index into the kernels' signature tables; the order of SIGNATURES
enum { SIG_NEWTONIAN = 0, SIG_SPECTROGRAPH = 1, SIG_SPECTROGRAPH_TILTED = 2, NUM_SIGNATURES = 3 };

constexpr Signature SIGNATURES[NUM_SIGNATURES] = {
    {0, 4, {OPTK_SIGNATURE_NEWTONIAN}},
    {FEAT_ALL, 5, {OPTK_SIGNATURE_SPECTROGRAPH}},
    {FEAT_ALL, 5, {OPTK_SIGNATURE_SPECTROGRAPH_TILTED}},
};

// the signature a system's records match (features: its kernel variant),
// or -1
inline int find_signature(const optk_surface *surfaces, int num_surfaces, int features) {
  for (int k = 0; k < NUM_SIGNATURES; ++k) {
    const Signature &sig = SIGNATURES[k];
    if (sig.num_surfaces != num_surfaces || sig.features != features) continue;
    bool same = true;
    for (int i = 0; i < num_surfaces && same; ++i) same = kinds_of(surfaces[i]) == sig.kinds[i];
    if (same) return k;
  }
  return -1;
}

}  // namespace optk
