// pbrt_scene.h -- reads pbrt-v2 scene files (.pbrt) into what the ray-space accelerators and
// the breadth-first renderer loop take.
//
// The reader follows pbrt-v2's scene-description front end for the parts that reach the
// accelerators: the lexer (src/core/pbrtlex.ll), the grammar and parameter typing
// (src/core/pbrtparse.yy:160-810) and the API state machine (src/core/api.cpp:283-1110).
//   - Triangles: Shape "trianglemesh" (src/shapes/trianglemesh.cpp:363-440), "heightfield"
//     (src/shapes/heightfield.cpp:62-116) and "loopsubdiv" (Loop subdivision to its limit
//     surface, src/shapes/loopsubdiv.cpp:148-502).  Vertices are taken to world space by the current
//     transform as TriangleMesh's constructor does, with pbrt's FP32 operation order.  Meshes
//     are kept in directive order: the order in which the accelerators flatten them
//     (rayhierarchy.cpp:157-199).
//   - The transform stack: Identity, Translate, Scale, Rotate, LookAt, Transform,
//     ConcatTransform, CoordinateSystem, CoordSysTransform, ActiveTransform, TransformTimes,
//     AttributeBegin/End, TransformBegin/End (api.cpp:704-910, transform.cpp:137-260).
//   - Accelerator name and parameters, for MakeAccelerator (api.cpp:576-596, 820-825).
//   - The renderer settings: Camera "perspective" (fov, halffov, frameaspectratio,
//     screenwindow), Film xresolution / yresolution, Sampler samples per pixel, LightSource
//     "point" (from, I, scale), SurfaceIntegrator directlighting / path maxdepth.
//
// Geometry that the ray-space accelerators never receive is parsed and dropped, as in the
// reference, where only GeometricPrimitives holding a Triangle are flattened
// (rayhierarchy.cpp:164-168):
//   - non-triangle shapes (sphere, disk, cylinder, cone, paraboloid, hyperboloid);
//   - shapes inside ObjectBegin / ObjectEnd, and ObjectInstance (TransformedPrimitives);
//   - shapes under an animated transform (TransformedPrimitive, api.cpp:1010-1040).
// "nurbs" refines into a triangle mesh in pbrt but is not provided here: it is reported and
// skipped.  Materials, textures, volumes, filters and area lights do not change the traced
// geometry and are ignored.
//
// Error behaviour follows pbrt's: a syntax error or an unreadable file is Severe(); what pbrt
// reports with Error() or Warning() and then ignores is recorded in SceneFile::messages.
#pragma once
#include <string>
#include <vector>

#include "breadth_first.h"
#include "pbrt_accel.h"

namespace pbrt_amd {

struct SceneFile {
  std::vector<TriangleMesh> meshes;  // world space, directive order
  // RenderOptions defaults (api.cpp:193-206)
  std::string accelerator = "bvh";
  ParamSet acceleratorParams;
  std::string surfaceIntegrator = "directlighting";
  std::string renderer = "sampler";
  std::string camera = "perspective";
  std::string sampler = "lowdiscrepancy";
  RenderSettings render;  // film, samples, camera, point light and path depth
  bool hasPointLight = false;
  int skippedShapes = 0;              // shapes pbrt keeps but the accelerators never see
  std::vector<std::string> messages;  // pbrt's Error / Warning output, "file:line: text"
};

// Parses the scene file at `path`; Include names resolve against the directory of that file,
// as pbrt's search directory (fileutil.cpp ResolveFilename).
void ParseSceneFile(const std::string& path, SceneFile* out);
// Parses scene text; `dir` is the directory Include names resolve against.
void ParseSceneString(const std::string& text, const std::string& dir, SceneFile* out);

// Builds the scene's accelerator with MakeAccelerator (or `accelerator_override` when it is
// not empty; the scene's parameters go with it only when the names match) and renders the
// scene with RenderBreadthFirst.  The accelerator must be one of "naive", "rayhierarchy" or
// "ray-bvh".
RenderStats RenderScene(const SceneFile& scene, const std::string& accelerator_override,
                        std::vector<float>* rgb);

}  // namespace pbrt_amd
