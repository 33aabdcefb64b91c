/* spectrapack_validate.h — C ABI of the exact host validation of final
 * placements (libspectrapack_compute.so; host code only, no GPU needed).
 *
 * Reference interface replaced (paths relative to the reference repository):
 *   engine/pack_geometry/src/validation.cpp:517-634 validate /
 *   validate_incremental (pair solids, then containment), with the verdict
 *   codes of that file.  The geometry is the exact image of the part mesh under
 *   each pose (validation_kernel.cpp:490-517 rotation H(q)/|q|^2).
 */
#ifndef SPECTRAPACK_VALIDATE_H
#define SPECTRAPACK_VALIDATE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double translation_mm[3];
    double rotation_xyzw[4];
} sp_pose;

enum sp_validity { SP_VALID = 0, SP_INVALID = 1, SP_INDETERMINATE = 2 };

typedef struct {
    int32_t validity;             /* sp_validity */
    uint32_t reserved;
    uint64_t first, second;       /* affected copies (indices), UINT64_MAX when none */
    uint64_t pair_tests;          /* copy pairs (and copy-container pairs) classified by their surfaces */
    uint64_t triangle_tests;      /* triangle pairs classified */
    uint64_t exact_evaluations;   /* predicates decided by exact arithmetic */
    char code[48];                /* "VALID", "VALIDATION_PAIR", "VALIDATION_PAIR_UNRESOLVED",
                                     "VALIDATION_CONTAINER", "VALIDATION_CONTAINER_UNRESOLVED",
                                     "KERNEL_POSE_INVALID" */
} sp_validation_report;

/* The container is the box [0, box_dims] per axis when box_dims is not null,
 * else the closed mesh (container_vertices, container_triangles) whose
 * interior is the free region.  Part and container meshes are closed and
 * consistently oriented; vertices are x, y, z doubles, triangles three
 * indices.  Returns 0, or -1 for a null or out-of-range argument. */
int sp_validate_placements(const double* part_vertices, uint32_t part_vertex_count, const uint32_t* part_triangles,
                           uint32_t part_triangle_count, const sp_pose* poses, uint64_t pose_count,
                           const double* box_dims, const double* container_vertices, uint32_t container_vertex_count,
                           const uint32_t* container_triangles, uint32_t container_triangle_count,
                           double pair_clearance_mm, double wall_clearance_mm, sp_validation_report* out);

/* Incremental form for a placement loop's accept verdict: try_add classifies
 * one pose against the container and every copy kept so far and keeps it
 * only when it is valid.  The validator copies the meshes. */
typedef struct sp_validator sp_validator;
int sp_validator_create(const double* part_vertices, uint32_t part_vertex_count, const uint32_t* part_triangles,
                        uint32_t part_triangle_count, const double* box_dims, const double* container_vertices,
                        uint32_t container_vertex_count, const uint32_t* container_triangles,
                        uint32_t container_triangle_count, double pair_clearance_mm, double wall_clearance_mm,
                        sp_validator** out);
int sp_validator_try_add(sp_validator* v, const sp_pose* pose, sp_validation_report* out);
void sp_validator_destroy(sp_validator* v);

#ifdef __cplusplus
}
#endif

#endif
