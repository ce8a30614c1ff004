// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the declarations below follow its public headers)
// SPDX-License-Identifier: MIT
// box2d_amd public header: simulation constants.
//
// Same values as the reference (include/box2d/constants.h:14-73); the step results depend on them.
#pragma once

#include "math_functions.h"

#define B2_HUGE ( 1.0e5f * b2GetLengthUnitsPerMeter() )
#define B2_MAX_WORKERS 32
#define B2_MAX_TASKS 256
#define B2_GRAPH_COLOR_COUNT 24
#define B2_LINEAR_SLOP ( 0.005f * b2GetLengthUnitsPerMeter() )
#ifndef B2_MAX_WORLDS
#define B2_MAX_WORLDS 4096
#endif
#ifndef B2_NAME_LENGTH
#define B2_NAME_LENGTH 10
#endif
#define B2_MAX_ROTATION ( 0.25f * B2_PI )
#define B2_SPECULATIVE_DISTANCE ( 4.0f * B2_LINEAR_SLOP )
#define B2_CONTACT_RECYCLE_DISTANCE ( 10.0f * B2_LINEAR_SLOP )
#define B2_CONTACT_RECYCLE_COS_ANGLE ( 0.98f )
#define B2_MAX_AABB_MARGIN ( 0.05f * b2GetLengthUnitsPerMeter() )
#define B2_AABB_MARGIN_FRACTION 0.125f
#define B2_TIME_TO_SLEEP 0.5f
