// ORACLE (test infrastructure only; never linked into the product library).
//
// C-ABI of the CPU restatement of the reference's coupled XPBD substep
// (soft_body.cpp:1021-1460 with suspension.cpp:2072-2610), run in the
// reference's coloured order (workerThreads >= 1) on one thread. Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
#pragma once

#include "../include/simwing_xpbd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_body oracle_body;

/* Builds the body from a definition; hinges are inserted in the build's
 * colour-major order (oracle_colouring.hpp), as the fixture harness does. */
int oracle_create(const sw_definition* definition, oracle_body** out);
void oracle_destroy(oracle_body* body);
int oracle_set_face_pressure(oracle_body* body, const double* pressure, size_t count);
int oracle_set_external_forces(oracle_body* body, const double* forces, size_t node_count);
/* One Structure::step: replays the external loads for every substep. */
int oracle_step(oracle_body* body, const sw_step_settings* settings);
/* State in definition order; segment lambdas in string-id order; any
 * pointer may be NULL. payload: com, q(w,x,y,z), v, w = 13 doubles. */
int oracle_frame_fields(const oracle_body* body, double* vertex_position, double* vertex_velocity,
                        double* pending_force, double* line_length, double* line_violation);
int oracle_state(const oracle_body* body, double* position, double* previous, double* velocity,
                 double* constraint_lambda, double* membrane_lambda, double* membrane_resultant,
                 double* membrane_residual, double* dihedral_lambda, double* segment_lambda, double* payload,
                 double* suspension_residual);
/* Colour orders (family 0 distance, 1 membrane, 2 dihedral). */
int oracle_colouring(const oracle_body* body, int family, uint64_t* order, size_t* order_count,
                     uint64_t* offsets, size_t* offset_count);
const char* oracle_last_error(void);

/* Test hook: perturb every dihedral angle by `ulps` units in the last place
 * (used only to measure how far a one-ulp difference grows). */
void oracle_set_angle_perturbation(oracle_body* body, int ulps);

/* Transcendentals of the body's steps: 0 = the C library (as the reference),
 * 1 = the portable double-double functions the device kernels use. */
void oracle_set_portable_math(oracle_body* body, int portable);

/* Evaluates atan2(a, b) (function 2), sin(a) (4), cos(a) (5) or acos(a) (6)
 * with the C library (portable = 0) or the portable functions (1), and the
 * same functions in quadruple precision rounded to double. */
void oracle_math(int portable, int function, const double* a, const double* b, double* out, size_t count);
void oracle_quad_math(int function, const double* a, const double* b, double* out, size_t count);

/* Tangent build only (liboracle_xpbd_tan.so; other builds report
 * SW_LOGIC_ERROR): seeds one direction of the parameters (per membrane
 * log-compliance, compliance C e^theta; per segment rest length, definition
 * order; per triangle pressure; NULL = zero) after the face pressure is set,
 * and reads the tangent of every node position (3 per node) after steps. */
int oracle_set_tangent(oracle_body* body, const double* d_log_compliance, const double* d_rest,
                       const double* d_pressure);
int oracle_tangent_state(const oracle_body* body, double* d_position);

#ifdef __cplusplus
}
#endif
