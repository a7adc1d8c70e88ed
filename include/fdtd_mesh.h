/* fdtd_mesh.h - automatic nonuniform mesh (MeshMode::AUTO_NONUNIFORM).
 *
 * C ABI for the reference's AutoMeshGenerator (src/auto_mesh_generator.hpp),
 * restated on the host: the same grid lines, bit for bit, for the same
 * configuration.  A caller builds the spacing arrays of an auto-meshed
 * scene from it and passes them, with the bounds, in fdtd_scene_spec
 * (dx/dy/dz and x_bounds/y_bounds/z_bounds).
 *
 * Errors: FDTD_EINVAL (from fdtd_hip.h) for a bad argument, and when
 * `capacity` is smaller than the axis's line count (*n_lines then says how
 * many are needed; lines == NULL with capacity 0 asks for the count).
 */
#ifndef FDTD_MESH_H
#define FDTD_MESH_H

#ifdef __cplusplus
extern "C" {
#endif

/* AutoMeshConfig (src/auto_mesh_generator.hpp:73-117). */
typedef struct {
    double lambda_min;          /* m; replaced by c0 / f_max when f_max > 0 */
    double f_max;               /* Hz, or 0 */
    int    mesh_accuracy;       /* 1 -> 6 ppw, 2 -> 10, 3 -> 14, then +4 per level */
    double ppw_override;        /* > 0: points per wavelength, overriding the accuracy */
    double dx_min, dx_max;      /* m, clamp of every cell size */
    double max_grading_ratio;   /* adjacent-cell ratio limit, sqrt(2) by default */
    int    pml_graded;          /* PMLStrategy: 0 UNIFORM (boundary cell repeated), 1 GRADED */
    int    override_enabled;    /* mesh_override_enabled */
} fdtd_auto_mesh_config;

/* AutoMeshStructureBounds (src/auto_mesh_generator.hpp:120-125): an
 * axis-aligned box in physical coordinates with its refractive index. */
typedef struct {
    double x_min, x_max, y_min, y_max, z_min, z_max;
    double n;
} fdtd_mesh_bounds;

/* MeshOverrideRegion (src/auto_mesh_generator.hpp:53-63), physical coordinates. */
typedef struct {
    int    enabled;
    double x0, x1, y0, y1, z0, z1;
    double dx, dy, dz;
    int    priority;            /* carried, unused by the generator (as in the reference) */
} fdtd_mesh_override;

/* The total-domain grid lines of one axis (0 = x, 1 = y, 2 = z):
 * AutoMeshGenerator::generate (src/auto_mesh_generator.hpp:156-253) for that
 * axis, i.e. generate_physical_mesh_1d (:256-276) then extend_with_pml_1d
 * (:279-375).  `domain` is the physical domain size along the axis, `npml`
 * the CPML cells added on each side; the lines start at 0, the physical
 * domain spans lines[npml] .. lines[n_lines - 1 - npml], and cell i has
 * width lines[i + 1] - lines[i] (how generate() fills GridSpacing::dx).
 * Replaces the reference's generator.generate(...) call in make_context
 * (src/params.hpp:209-211). */
int fdtd_auto_mesh_axis(const fdtd_auto_mesh_config* cfg, int axis, double domain, int npml,
                        int n_structures, const fdtd_mesh_bounds* structures,
                        int n_overrides, const fdtd_mesh_override* overrides,
                        double background_n, double* lines, long capacity, long* n_lines);

/* The mesh box of a structure (type FDTD_SHAPE_*, params and index as in
 * fdtd_structure_def of fdtd_scene.h): register_structure_bounds_for_auto_mesh
 * (src/params.hpp:540-553) - a box as given, a sphere's and a z-cylinder's
 * bounding box. */
int fdtd_structure_mesh_bounds(int type, const double* params, double n, fdtd_mesh_bounds* out);

#ifdef __cplusplus
}
#endif

#endif /* FDTD_MESH_H */
