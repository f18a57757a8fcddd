// Joint rows and launch declarations (joints.hip).
#pragma once

#include "rp3d_internal.h"

namespace rp3dg {

// joint kinds in solve order (ConstraintSolverSystem.cpp:94-113); values are internal
constexpr int JT_BALLSOCKET = 0;
constexpr int JT_FIXED = 1;
constexpr int JT_HINGE = 2;
constexpr int JT_SLIDER = 3;
constexpr int JT_KINDS = 4;

constexpr int JF_LIMIT = 1;
constexpr int JF_MOTOR = 2;
constexpr int JF_LOWER_VIOLATED = 4;
constexpr int JF_UPPER_VIOLATED = 8;   // ball-and-socket: JF_LOWER_VIOLATED marks the violated cone limit
// JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS (configuration.h:101): the position error enters the
// velocity rows as a bias of 0.2 / dt times the error, and the position solver skips the joint
constexpr int JF_BAUMGARTE = 16;

// One entry per joint, sorted by (kind, colour, creation index).
struct DevJoints {
    int n;
    int* type;
    int2* bodies;
    float4 *la1, *la2;            // local anchor points (body space)
    float4 *ax1, *ax2;            // hinge axis in body spaces
    float4* qinit;                // initial orientation difference inverse (fixed, hinge)
    int* flags;                   // JF_* (violation bits persist between steps)
    float4* lim;                  // lower, upper, motor speed, max motor torque / force
                                  // (ball-and-socket: x = cone half angle)
    float4 *impT, *impR, *impL;   // accumulated impulses: translation, rotation, (lower, upper, motor)
                                  // (ball-and-socket: impL.x = cone limit; slider: impT.xy = 2 rows)
    // per-step rows
    float4 *i1r0, *i1r1, *i1r2, *i2r0, *i2r1, *i2r2;
    float4 *r1, *r2;
    float4 *kt0, *kt1, *kt2;      // inverse translation mass matrix (slider: 2x2 in kt0)
    float4 *kr0, *kr1, *kr2;      // inverse rotation mass matrix (fixed, slider: 3x3, hinge: 2x2 in kr0)
    float4 *a1, *b2xa1, *c2xa1;   // hinge axes; ball-socket: cone axis in a1; slider: axis, n1, n2
    float* imlm;                  // inverse mass of the hinge limit / motor row, cone limit, slider limit
    float* imm;                   // slider: inverse mass of the motor row
    // slider: (r1 + u) x {n1, n2, axis} and r2 x {n1, n2, axis}; r2 x axis persists between steps
    // (the limit row's inverse mass of the next step reads the value the position solver left)
    float4 *un1, *un2, *ua, *rn1, *rn2, *ra;
    // Baumgarte biases of the step (JF_BAUMGARTE joints only):
    //   ball-and-socket: bias0 = (translation bias, cone limit bias)
    //   fixed:           bias0.xyz = translation bias, bias1.xyz = rotation bias
    //   hinge:           bias0.xyz = translation bias, bias1 = (rotation bias x, y, lower limit, upper limit)
    //   slider:          bias0 = (translation bias x, y, lower limit, upper limit), bias1.xyz = rotation bias
    float4 *bias0, *bias1;
    // SolveSliderJointSystem::solvePositionConstraint returns (SolveSliderJointSystem.cpp:527) at the first
    // Baumgarte slider it meets, so a world's sliders from that one on get no position correction:
    // srank = rank of a slider in its world's solve order, bfirst[world] = smallest srank of an enabled
    // Baumgarte slider of the step (INT_MAX: none).  bfirst is null when no world has a Baumgarte slider.
    int* srank;
    int* bfirst;
};

__global__ void jt_init(DevBodies B, DevJoints J, float dt);
__global__ void jt_warmstart(DevBodies B, DevJoints J, const int* range, int slot);
__global__ void jt_solve_velocity(DevBodies B, DevJoints J, const int* range, int slot, float dt);
__global__ void jt_solve_position(DevBodies B, DevJoints J, const int* range, int slot);

}  // namespace rp3dg
