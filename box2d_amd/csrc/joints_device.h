// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the joint solve arithmetic below restates its
// src/*_joint.c in the same operation order)
// SPDX-License-Identifier: MIT
// GPU joint solver: warm start and solve of every joint type (revolute, weld, prismatic, wheel,
// distance, motor, mover, pogo; filter joints only take part in collision filtering and islands).
// Each function is one lane's work on one joint and keeps the reference's operation order, so its
// floats are the reference's:
//   revolute  src/revolute_joint.c:217-500
//   weld      src/weld_joint.c:148-452 (the non-block solve, B2_WELD_BLOCK_SOLVE 0)
//   prismatic src/prismatic_joint.c:295-664
//   wheel     src/wheel_joint.c:197-520
//   distance  src/distance_joint.c:250-541
//   motor     src/motor_joint.c:179-430
//   mover     src/mover_joint.c:47-170
//   pogo      src/pogo_joint.c:98-280
// Bodies is a body view of solver_device.h (device memory by body id, or LDS by slot).
#pragma once

namespace b2g
{

// The device path runs with one length unit per meter (world creation rejects other settings).
constexpr float kLengthUnitsPerMeter = 1.0f;

// Warm-start impulses of one joint: body A receives v -= mA * P, w -= iA * LA and body B
// v += mB * P, w += iB * LB. They depend on the accumulated impulses, the joint frames and the two
// bodies' delta positions and rotations, never on velocities. Returns false for joints with no
// warm start (filter, and pogo joints without a spring).
B2G_DEVI bool JointWarmStartTerms( const JointSim* j, b2Rot dqA, b2Vec2 dpA, b2Rot dqB, b2Vec2 dpB, b2Vec2& P, float& LA,
								   float& LB )
{
	switch ( j->type )
	{
		case b2_revoluteJoint:
		{
			b2Vec2 rA = b2RotateVector( dqA, j->frameA.p );
			b2Vec2 rB = b2RotateVector( dqB, j->frameB.p );
			float axialImpulse = j->springImpulse + j->motorImpulse + j->lowerImpulse - j->upperImpulse;
			P = j->linearImpulse;
			LA = b2Cross( rA, j->linearImpulse ) + axialImpulse;
			LB = b2Cross( rB, j->linearImpulse ) + axialImpulse;
			return true;
		}
		case b2_weldJoint:
		{
			b2Vec2 rA = b2RotateVector( dqA, j->frameA.p );
			b2Vec2 rB = b2RotateVector( dqB, j->frameB.p );
			P = j->linearImpulse;
			LA = b2Cross( rA, j->linearImpulse ) + j->angularImpulse;
			LB = b2Cross( rB, j->linearImpulse ) + j->angularImpulse;
			return true;
		}
		case b2_motorJoint:
		{
			b2Vec2 rA = b2RotateVector( dqA, j->frameA.p );
			b2Vec2 rB = b2RotateVector( dqB, j->frameB.p );
			b2Vec2 linearImpulse = b2Add( j->linearVelocityImpulse, j->linearSpringImpulse );
			float angularImpulse = j->angularVelocityImpulse + j->angularSpringImpulse;
			P = linearImpulse;
			LA = b2Cross( rA, linearImpulse ) + angularImpulse;
			LB = b2Cross( rB, linearImpulse ) + angularImpulse;
			return true;
		}
		case b2_prismaticJoint:
		{
			b2Vec2 rA = b2RotateVector( dqA, j->frameA.p );
			b2Vec2 rB = b2RotateVector( dqB, j->frameB.p );
			b2Vec2 d = b2Add( b2Add( b2Sub( dpB, dpA ), j->deltaCenter ), b2Sub( rB, rA ) );
			b2Vec2 axisA = b2RotateVector( j->frameA.q, b2Vec2{ 1.0f, 0.0f } );
			axisA = b2RotateVector( dqA, axisA );

			// the impulse acts at body B's anchor
			float a1 = b2Cross( b2Add( rA, d ), axisA );
			float a2 = b2Cross( rB, axisA );
			float axialImpulse = j->springImpulse + j->motorImpulse + j->lowerImpulse - j->upperImpulse;

			b2Vec2 perpA = b2LeftPerp( axisA );
			float s1 = b2Cross( b2Add( rA, d ), perpA );
			float s2 = b2Cross( rB, perpA );
			float perpImpulse = j->pointImpulse.x;
			float angleImpulse = j->pointImpulse.y;

			P = b2Add( b2MulSV( axialImpulse, axisA ), b2MulSV( perpImpulse, perpA ) );
			LA = axialImpulse * a1 + perpImpulse * s1 + angleImpulse;
			LB = axialImpulse * a2 + perpImpulse * s2 + angleImpulse;
			return true;
		}
		case b2_wheelJoint:
		{
			b2Vec2 rA = b2RotateVector( dqA, j->frameA.p );
			b2Vec2 rB = b2RotateVector( dqB, j->frameB.p );
			b2Vec2 d = b2Add( b2Add( b2Sub( dpB, dpA ), j->deltaCenter ), b2Sub( rB, rA ) );
			b2Vec2 axisA = b2RotateVector( j->frameA.q, b2Vec2{ 1.0f, 0.0f } );
			axisA = b2RotateVector( dqA, axisA );
			b2Vec2 perpA = b2LeftPerp( axisA );

			float a1 = b2Cross( b2Add( d, rA ), axisA );
			float a2 = b2Cross( rB, axisA );
			float s1 = b2Cross( b2Add( d, rA ), perpA );
			float s2 = b2Cross( rB, perpA );

			float axialImpulse = j->springImpulse + j->lowerImpulse - j->upperImpulse;
			P = b2Add( b2MulSV( axialImpulse, axisA ), b2MulSV( j->perpImpulse, perpA ) );
			LA = axialImpulse * a1 + j->perpImpulse * s1 + j->motorImpulse;
			LB = axialImpulse * a2 + j->perpImpulse * s2 + j->motorImpulse;
			return true;
		}
		case b2_distanceJoint:
		{
			b2Vec2 rA = b2RotateVector( dqA, j->anchorA );
			b2Vec2 rB = b2RotateVector( dqB, j->anchorB );
			b2Vec2 ds = b2Add( b2Sub( dpB, dpA ), b2Sub( rB, rA ) );
			b2Vec2 separation = b2Add( j->deltaCenter, ds );
			b2Vec2 axis = b2Normalize( separation );

			float axialImpulse = j->distanceImpulse + j->lowerImpulse - j->upperImpulse + j->motorImpulse;
			P = b2MulSV( axialImpulse, axis );
			LA = b2Cross( rA, P );
			LB = b2Cross( rB, P );
			return true;
		}
		case b2_moverJoint:
		{
			// linear only; the mover's preparation zeroes its inverse inertias
			P = j->linearVelocityImpulse;
			LA = 0.0f;
			LB = 0.0f;
			return true;
		}
		case b2_pogoJoint:
		{
			// a pogo without spring has no warm start (its impulse is zeroed when prepared)
			if ( j->hertz == 0.0f )
			{
				return false;
			}
			b2Vec2 rA = b2RotateVector( dqA, j->frameA.p );
			b2Vec2 rB = b2RotateVector( dqB, j->frameB.p );
			P = b2MulSV( j->pogoImpulse, j->normal );
			LA = b2Cross( rA, P );
			LB = b2Cross( rB, P );
			return true;
		}
		default:
			return false;
	}
}

template <typename Bodies>
B2G_DEVI void WarmStartJoint( JointSim* joint, const Bodies& B )
{
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 P;
	float LA, LB;
	if ( JointWarmStartTerms( joint, sA.deltaRotation, sA.deltaPosition, sB.deltaRotation, sB.deltaPosition, P, LA, LB ) == false )
	{
		return;
	}

	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = b2MulSub( sA.linearVelocity, mA, P );
		sA.angularVelocity -= iA * LA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = b2MulAdd( sB.linearVelocity, mB, P );
		sB.angularVelocity += iB * LB;
	}

	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// Soft limit row shared by the prismatic, wheel, distance and revolute limits: speculative when
// C > 0, soft (useBias) or rigid otherwise.
struct LimitRow
{
	float bias;
	float massScale;
	float impulseScale;
};

B2G_DEVI LimitRow MakeLimitRow( float C, float speculativeBias, bool useBias, const Softness& softness )
{
	LimitRow r = { 0.0f, 1.0f, 0.0f };
	if ( C > 0.0f )
	{
		r.bias = speculativeBias;
	}
	else if ( useBias )
	{
		r.bias = softness.biasRate * C;
		r.massScale = softness.massScale;
		r.impulseScale = softness.impulseScale;
	}
	return r;
}

// ---------------------------------------------------------------------------------------------
// Revolute

template <typename Bodies>
B2G_DEVI void SolveRevolute( JointSim* joint, const Bodies& B, bool useBias, float h, float inv_h )
{
	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );

	BodyState sA = ToState( B.Load( ia ) );
	BodyState* stateA = &sA;
	BodyState sB = ToState( B.Load( ib ) );
	BodyState* stateB = &sB;

	b2Vec2 vA = stateA->linearVelocity;
	float wA = stateA->angularVelocity;
	b2Vec2 vB = stateB->linearVelocity;
	float wB = stateB->angularVelocity;

	b2Rot qA = b2MulRot( stateA->deltaRotation, joint->frameA.q );
	b2Rot qB = b2MulRot( stateB->deltaRotation, joint->frameB.q );
	b2Rot relQ = b2InvMulRot( qA, qB );

	bool fixedRotation = ( iA + iB == 0.0f );

	if ( joint->enableSpring && fixedRotation == false )
	{
		float jointAngle = Atan2( relQ.s, relQ.c );
		float jointAngleDelta = b2UnwindAngle( jointAngle - joint->targetAngle );
		float C = jointAngleDelta;
		float bias = joint->springSoftness.biasRate * C;
		float massScale = joint->springSoftness.massScale;
		float impulseScale = joint->springSoftness.impulseScale;
		float Cdot = wB - wA;
		float impulse = -massScale * joint->axialMass * ( Cdot + bias ) - impulseScale * joint->springImpulse;
		joint->springImpulse += impulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}

	if ( joint->enableMotor && fixedRotation == false )
	{
		float Cdot = wB - wA - joint->motorSpeed;
		float impulse = -joint->axialMass * Cdot;
		float oldImpulse = joint->motorImpulse;
		float maxImpulse = h * joint->maxMotorTorque;
		joint->motorImpulse = b2ClampFloat( joint->motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint->motorImpulse - oldImpulse;
		wA -= iA * impulse;
		wB += iB * impulse;
	}

	if ( joint->enableLimit && fixedRotation == false )
	{
		float jointAngle = Atan2( relQ.s, relQ.c );

		{
			float C = jointAngle - joint->lowerAngle;
			LimitRow r = MakeLimitRow( C, C * inv_h, useBias, joint->constraintSoftness );
			float Cdot = wB - wA;
			float oldImpulse = joint->lowerImpulse;
			float impulse = -r.massScale * joint->axialMass * ( Cdot + r.bias ) - r.impulseScale * oldImpulse;
			joint->lowerImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint->lowerImpulse - oldImpulse;
			wA -= iA * impulse;
			wB += iB * impulse;
		}

		{
			float C = joint->upperAngle - jointAngle;
			LimitRow r = MakeLimitRow( C, C * inv_h, useBias, joint->constraintSoftness );
			float Cdot = wA - wB;
			float oldImpulse = joint->upperImpulse;
			float impulse = -r.massScale * joint->axialMass * ( Cdot + r.bias ) - r.impulseScale * oldImpulse;
			joint->upperImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint->upperImpulse - oldImpulse;
			wA += iA * impulse;
			wB -= iB * impulse;
		}
	}

	{
		b2Vec2 rA = b2RotateVector( stateA->deltaRotation, joint->frameA.p );
		b2Vec2 rB = b2RotateVector( stateB->deltaRotation, joint->frameB.p );

		b2Vec2 Cdot = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );

		b2Vec2 bias = { 0.0f, 0.0f };
		float massScale = 1.0f;
		float impulseScale = 0.0f;
		if ( useBias )
		{
			b2Vec2 dcA = stateA->deltaPosition;
			b2Vec2 dcB = stateB->deltaPosition;
			b2Vec2 separation = b2Add( b2Add( b2Sub( dcB, dcA ), b2Sub( rB, rA ) ), joint->deltaCenter );
			bias = b2MulSV( joint->constraintSoftness.biasRate, separation );
			massScale = joint->constraintSoftness.massScale;
			impulseScale = joint->constraintSoftness.impulseScale;
		}

		b2Mat22 K;
		K.cx.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
		K.cy.x = -rA.y * rA.x * iA - rB.y * rB.x * iB;
		K.cx.y = K.cy.x;
		K.cy.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
		b2Vec2 b = b2Solve22( K, b2Add( Cdot, bias ) );

		b2Vec2 impulse;
		impulse.x = -massScale * b.x - impulseScale * joint->linearImpulse.x;
		impulse.y = -massScale * b.y - impulseScale * joint->linearImpulse.y;
		joint->linearImpulse.x += impulse.x;
		joint->linearImpulse.y += impulse.y;

		vA = b2MulSub( vA, mA, impulse );
		wA -= iA * b2Cross( rA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}

	if ( stateA->flags & kDynamicFlag )
	{
		stateA->linearVelocity = vA;
		stateA->angularVelocity = wA;
	}
	if ( stateB->flags & kDynamicFlag )
	{
		stateB->linearVelocity = vB;
		stateB->angularVelocity = wB;
	}

	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Weld

template <typename Bodies>
B2G_DEVI void SolveWeld( JointSim* joint, const Bodies& B, bool useBias )
{
	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 vA = sA.linearVelocity;
	float wA = sA.angularVelocity;
	b2Vec2 vB = sB.linearVelocity;
	float wB = sB.angularVelocity;

	// angular constraint
	{
		b2Rot qA = b2MulRot( sA.deltaRotation, joint->frameA.q );
		b2Rot qB = b2MulRot( sB.deltaRotation, joint->frameB.q );
		b2Rot relQ = b2InvMulRot( qA, qB );
		float jointAngle = Atan2( relQ.s, relQ.c );

		float bias = 0.0f;
		float massScale = 1.0f;
		float impulseScale = 0.0f;
		if ( useBias || joint->angularHertz > 0.0f )
		{
			float C = jointAngle;
			bias = joint->angularSpring.biasRate * C;
			massScale = joint->angularSpring.massScale;
			impulseScale = joint->angularSpring.impulseScale;
		}

		float Cdot = wB - wA;
		float impulse = -massScale * joint->axialMass * ( Cdot + bias ) - impulseScale * joint->angularImpulse;
		joint->angularImpulse += impulse;

		wA -= iA * impulse;
		wB += iB * impulse;
	}

	// linear constraint
	{
		b2Vec2 rA = b2RotateVector( sA.deltaRotation, joint->frameA.p );
		b2Vec2 rB = b2RotateVector( sB.deltaRotation, joint->frameB.p );

		b2Vec2 bias = { 0.0f, 0.0f };
		float massScale = 1.0f;
		float impulseScale = 0.0f;
		if ( useBias || joint->linearHertz > 0.0f )
		{
			b2Vec2 dcA = sA.deltaPosition;
			b2Vec2 dcB = sB.deltaPosition;
			b2Vec2 C = b2Add( b2Add( b2Sub( dcB, dcA ), b2Sub( rB, rA ) ), joint->deltaCenter );

			bias = b2MulSV( joint->linearSpring.biasRate, C );
			massScale = joint->linearSpring.massScale;
			impulseScale = joint->linearSpring.impulseScale;
		}

		b2Vec2 Cdot = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );

		b2Mat22 K;
		K.cx.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
		K.cy.x = -rA.y * rA.x * iA - rB.y * rB.x * iB;
		K.cx.y = K.cy.x;
		K.cy.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
		b2Vec2 b = b2Solve22( K, b2Add( Cdot, bias ) );

		b2Vec2 impulse = {
			-massScale * b.x - impulseScale * joint->linearImpulse.x,
			-massScale * b.y - impulseScale * joint->linearImpulse.y,
		};

		joint->linearImpulse = b2Add( joint->linearImpulse, impulse );

		vA = b2MulSub( vA, mA, impulse );
		wA -= iA * b2Cross( rA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}

	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = vA;
		sA.angularVelocity = wA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = vB;
		sB.angularVelocity = wB;
	}
	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Prismatic

template <typename Bodies>
B2G_DEVI void SolvePrismatic( JointSim* joint, const Bodies& B, bool useBias, float h, float inv_h )
{
	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 vA = sA.linearVelocity;
	float wA = sA.angularVelocity;
	b2Vec2 vB = sB.linearVelocity;
	float wB = sB.angularVelocity;

	b2Rot qA = b2MulRot( sA.deltaRotation, joint->frameA.q );
	b2Rot qB = b2MulRot( sB.deltaRotation, joint->frameB.q );
	b2Rot relQ = b2InvMulRot( qA, qB );

	// current anchors
	b2Vec2 rA = b2RotateVector( sA.deltaRotation, joint->frameA.p );
	b2Vec2 rB = b2RotateVector( sB.deltaRotation, joint->frameB.p );

	b2Vec2 d = b2Add( b2Add( b2Sub( sB.deltaPosition, sA.deltaPosition ), joint->deltaCenter ), b2Sub( rB, rA ) );

	b2Vec2 axisA = b2RotateVector( joint->frameA.q, b2Vec2{ 1.0f, 0.0f } );
	axisA = b2RotateVector( sA.deltaRotation, axisA );
	float translation = b2Dot( axisA, d );

	// torque arms of the axial force
	float a1 = b2Cross( b2Add( rA, d ), axisA );
	float a2 = b2Cross( rB, axisA );

	float k = mA + mB + iA * a1 * a1 + iB * a2 * a2;
	float axialMass = k > 0.0f ? 1.0f / k : 0.0f;

	Softness softness = joint->constraintSoftness;

	// spring (a real spring: applied during relax too)
	if ( joint->enableSpring )
	{
		float C = translation - joint->targetTranslation;
		float bias = joint->springSoftness.biasRate * C;
		float massScale = joint->springSoftness.massScale;
		float impulseScale = joint->springSoftness.impulseScale;

		float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
		float deltaImpulse = -massScale * axialMass * ( Cdot + bias ) - impulseScale * joint->springImpulse;
		joint->springImpulse += deltaImpulse;

		b2Vec2 P = b2MulSV( deltaImpulse, axisA );
		float LA = deltaImpulse * a1;
		float LB = deltaImpulse * a2;

		vA = b2MulSub( vA, mA, P );
		wA -= iA * LA;
		vB = b2MulAdd( vB, mB, P );
		wB += iB * LB;
	}

	if ( joint->enableMotor )
	{
		float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
		float impulse = axialMass * ( joint->motorSpeed - Cdot );
		float oldImpulse = joint->motorImpulse;
		float maxImpulse = h * joint->maxMotorForce;
		joint->motorImpulse = b2ClampFloat( joint->motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint->motorImpulse - oldImpulse;

		b2Vec2 P = b2MulSV( impulse, axisA );
		float LA = impulse * a1;
		float LB = impulse * a2;

		vA = b2MulSub( vA, mA, P );
		wA -= iA * LA;
		vB = b2MulAdd( vB, mB, P );
		wB += iB * LB;
	}

	if ( joint->enableLimit )
	{
		// the speculative distance is clamped to a quarter of the limit range
		float speculativeDistance = 0.25f * ( joint->upperTranslation - joint->lowerTranslation );

		// lower limit
		{
			float C = translation - joint->lowerTranslation;
			if ( C < speculativeDistance )
			{
				LimitRow r = MakeLimitRow( C, b2MinFloat( C, kLengthUnitsPerMeter ) * inv_h, useBias, softness );
				float oldImpulse = joint->lowerImpulse;
				float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
				float deltaImpulse = -axialMass * r.massScale * ( Cdot + r.bias ) - r.impulseScale * oldImpulse;
				joint->lowerImpulse = b2MaxFloat( oldImpulse + deltaImpulse, 0.0f );
				deltaImpulse = joint->lowerImpulse - oldImpulse;

				b2Vec2 P = b2MulSV( deltaImpulse, axisA );
				float LA = deltaImpulse * a1;
				float LB = deltaImpulse * a2;

				vA = b2MulSub( vA, mA, P );
				wA -= iA * LA;
				vB = b2MulAdd( vB, mB, P );
				wB += iB * LB;
			}
			else
			{
				joint->lowerImpulse = 0.0f;
			}
		}

		// upper limit, signs flipped so C and the impulse stay positive when active
		{
			float C = joint->upperTranslation - translation;
			if ( C < speculativeDistance )
			{
				LimitRow r = MakeLimitRow( C, b2MinFloat( C, kLengthUnitsPerMeter ) * inv_h, useBias, softness );
				float oldImpulse = joint->upperImpulse;
				float Cdot = b2Dot( axisA, b2Sub( vA, vB ) ) + a1 * wA - a2 * wB;
				float deltaImpulse = -axialMass * r.massScale * ( Cdot + r.bias ) - r.impulseScale * oldImpulse;
				joint->upperImpulse = b2MaxFloat( oldImpulse + deltaImpulse, 0.0f );
				deltaImpulse = joint->upperImpulse - oldImpulse;

				b2Vec2 P = b2MulSV( deltaImpulse, axisA );
				float LA = deltaImpulse * a1;
				float LB = deltaImpulse * a2;

				vA = b2MulAdd( vA, mA, P );
				wA += iA * LA;
				vB = b2MulSub( vB, mB, P );
				wB -= iB * LB;
			}
			else
			{
				joint->upperImpulse = 0.0f;
			}
		}
	}

	// point and angle constraint in block form
	{
		b2Vec2 perpA = b2LeftPerp( axisA );

		// torque arms of the perpendicular force
		float s1 = b2Cross( b2Add( d, rA ), perpA );
		float s2 = b2Cross( rB, perpA );

		b2Vec2 Cdot;
		Cdot.x = b2Dot( perpA, b2Sub( vB, vA ) ) + s2 * wB - s1 * wA;
		Cdot.y = wB - wA;

		b2Vec2 bias = { 0.0f, 0.0f };
		float massScale = 1.0f;
		float impulseScale = 0.0f;
		if ( useBias )
		{
			b2Vec2 C;
			C.x = b2Dot( perpA, d );
			C.y = Atan2( relQ.s, relQ.c );

			bias = b2MulSV( softness.biasRate, C );
			massScale = softness.massScale;
			impulseScale = softness.impulseScale;
		}

		float k11 = mA + mB + iA * s1 * s1 + iB * s2 * s2;
		float k12 = iA * s1 + iB * s2;
		float k22 = iA + iB;
		if ( k22 == 0.0f )
		{
			// bodies with fixed rotation
			k22 = 1.0f;
		}

		b2Mat22 K = { { k11, k12 }, { k12, k22 } };

		b2Vec2 b = b2Solve22( K, b2Add( Cdot, bias ) );
		b2Vec2 deltaImpulse;
		deltaImpulse.x = -massScale * b.x - impulseScale * joint->pointImpulse.x;
		deltaImpulse.y = -massScale * b.y - impulseScale * joint->pointImpulse.y;

		joint->pointImpulse.x += deltaImpulse.x;
		joint->pointImpulse.y += deltaImpulse.y;

		b2Vec2 P = b2MulSV( deltaImpulse.x, perpA );
		float LA = deltaImpulse.x * s1 + deltaImpulse.y;
		float LB = deltaImpulse.x * s2 + deltaImpulse.y;

		vA = b2MulSub( vA, mA, P );
		wA -= iA * LA;
		vB = b2MulAdd( vB, mB, P );
		wB += iB * LB;
	}

	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = vA;
		sA.angularVelocity = wA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = vB;
		sB.angularVelocity = wB;
	}
	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Wheel

template <typename Bodies>
B2G_DEVI void SolveWheel( JointSim* joint, const Bodies& B, bool useBias, float h, float inv_h )
{
	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 vA = sA.linearVelocity;
	float wA = sA.angularVelocity;
	b2Vec2 vB = sB.linearVelocity;
	float wB = sB.angularVelocity;

	bool fixedRotation = ( iA + iB == 0.0f );

	// current anchors
	b2Vec2 rA = b2RotateVector( sA.deltaRotation, joint->frameA.p );
	b2Vec2 rB = b2RotateVector( sB.deltaRotation, joint->frameB.p );

	b2Vec2 d = b2Add( b2Add( b2Sub( sB.deltaPosition, sA.deltaPosition ), joint->deltaCenter ), b2Sub( rB, rA ) );
	b2Vec2 axisA = b2RotateVector( joint->frameA.q, b2Vec2{ 1.0f, 0.0f } );
	axisA = b2RotateVector( sA.deltaRotation, axisA );
	float translation = b2Dot( axisA, d );

	float a1 = b2Cross( b2Add( d, rA ), axisA );
	float a2 = b2Cross( rB, axisA );

	// motor
	if ( joint->enableMotor && fixedRotation == false )
	{
		float Cdot = wB - wA - joint->motorSpeed;
		float impulse = -joint->motorMass * Cdot;
		float oldImpulse = joint->motorImpulse;
		float maxImpulse = h * joint->maxMotorTorque;
		joint->motorImpulse = b2ClampFloat( joint->motorImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint->motorImpulse - oldImpulse;

		wA -= iA * impulse;
		wB += iB * impulse;
	}

	// spring (a real spring: applied during relax too)
	if ( joint->enableSpring )
	{
		float C = translation;
		float bias = joint->springSoftness.biasRate * C;
		float massScale = joint->springSoftness.massScale;
		float impulseScale = joint->springSoftness.impulseScale;

		float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
		float impulse = -massScale * joint->axialMass * ( Cdot + bias ) - impulseScale * joint->springImpulse;
		joint->springImpulse += impulse;

		b2Vec2 P = b2MulSV( impulse, axisA );
		float LA = impulse * a1;
		float LB = impulse * a2;

		vA = b2MulSub( vA, mA, P );
		wA -= iA * LA;
		vB = b2MulAdd( vB, mB, P );
		wB += iB * LB;
	}

	if ( joint->enableLimit )
	{
		// lower limit
		{
			float C = translation - joint->lowerTranslation;
			LimitRow r = MakeLimitRow( C, C * inv_h, useBias, joint->constraintSoftness );

			float Cdot = b2Dot( axisA, b2Sub( vB, vA ) ) + a2 * wB - a1 * wA;
			float impulse = -r.massScale * joint->axialMass * ( Cdot + r.bias ) - r.impulseScale * joint->lowerImpulse;
			float oldImpulse = joint->lowerImpulse;
			joint->lowerImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint->lowerImpulse - oldImpulse;

			b2Vec2 P = b2MulSV( impulse, axisA );
			float LA = impulse * a1;
			float LB = impulse * a2;

			vA = b2MulSub( vA, mA, P );
			wA -= iA * LA;
			vB = b2MulAdd( vB, mB, P );
			wB += iB * LB;
		}

		// upper limit, signs flipped so C and the impulse stay positive when active
		{
			float C = joint->upperTranslation - translation;
			LimitRow r = MakeLimitRow( C, C * inv_h, useBias, joint->constraintSoftness );

			float Cdot = b2Dot( axisA, b2Sub( vA, vB ) ) + a1 * wA - a2 * wB;
			float impulse = -r.massScale * joint->axialMass * ( Cdot + r.bias ) - r.impulseScale * joint->upperImpulse;
			float oldImpulse = joint->upperImpulse;
			joint->upperImpulse = b2MaxFloat( oldImpulse + impulse, 0.0f );
			impulse = joint->upperImpulse - oldImpulse;

			b2Vec2 P = b2MulSV( impulse, axisA );
			float LA = impulse * a1;
			float LB = impulse * a2;

			vA = b2MulAdd( vA, mA, P );
			wA += iA * LA;
			vB = b2MulSub( vB, mB, P );
			wB -= iB * LB;
		}
	}

	// point to line
	{
		b2Vec2 perpA = b2LeftPerp( axisA );

		float bias = 0.0f;
		float massScale = 1.0f;
		float impulseScale = 0.0f;
		if ( useBias )
		{
			float C = b2Dot( perpA, d );
			bias = joint->constraintSoftness.biasRate * C;
			massScale = joint->constraintSoftness.massScale;
			impulseScale = joint->constraintSoftness.impulseScale;
		}

		float s1 = b2Cross( b2Add( d, rA ), perpA );
		float s2 = b2Cross( rB, perpA );
		float Cdot = b2Dot( perpA, b2Sub( vB, vA ) ) + s2 * wB - s1 * wA;

		float impulse = -massScale * joint->perpMass * ( Cdot + bias ) - impulseScale * joint->perpImpulse;
		joint->perpImpulse += impulse;

		b2Vec2 P = b2MulSV( impulse, perpA );
		float LA = impulse * s1;
		float LB = impulse * s2;

		vA = b2MulSub( vA, mA, P );
		wA -= iA * LA;
		vB = b2MulAdd( vB, mB, P );
		wB += iB * LB;
	}

	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = vA;
		sA.angularVelocity = wA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = vB;
		sB.angularVelocity = wB;
	}
	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Distance

template <typename Bodies>
B2G_DEVI void SolveDistance( JointSim* joint, const Bodies& B, bool useBias, float h, float inv_h )
{
	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 vA = sA.linearVelocity;
	float wA = sA.angularVelocity;
	b2Vec2 vB = sB.linearVelocity;
	float wB = sB.angularVelocity;

	// current anchors and separation
	b2Vec2 rA = b2RotateVector( sA.deltaRotation, joint->anchorA );
	b2Vec2 rB = b2RotateVector( sB.deltaRotation, joint->anchorB );
	b2Vec2 ds = b2Add( b2Sub( sB.deltaPosition, sA.deltaPosition ), b2Sub( rB, rA ) );
	b2Vec2 separation = b2Add( joint->deltaCenter, ds );

	float length = b2Length( separation );
	b2Vec2 axis = b2Normalize( separation );

	// soft when the spring is on and the limits do not pin the length
	if ( joint->enableSpring && ( joint->minLength < joint->maxLength || joint->enableLimit == 0 ) )
	{
		if ( joint->hertz > 0.0f )
		{
			b2Vec2 vr = b2Add( b2Sub( vB, vA ), b2Sub( b2CrossSV( wB, rB ), b2CrossSV( wA, rA ) ) );
			float cdot = b2Dot( axis, vr );
			float c = length - joint->length;
			float bias = joint->springSoftness.biasRate * c;

			float m = joint->springSoftness.massScale * joint->axialMass;
			float oldImpulse = joint->distanceImpulse;
			float impulse = -m * ( cdot + bias ) - joint->springSoftness.impulseScale * oldImpulse;

			joint->distanceImpulse = b2ClampFloat( joint->distanceImpulse + impulse, joint->lowerSpringForce * h, joint->upperSpringForce * h );
			impulse = joint->distanceImpulse - oldImpulse;

			b2Vec2 P = b2MulSV( impulse, axis );
			vA = b2MulSub( vA, mA, P );
			wA -= iA * b2Cross( rA, P );
			vB = b2MulAdd( vB, mB, P );
			wB += iB * b2Cross( rB, P );
		}

		if ( joint->enableMotor )
		{
			b2Vec2 vr = b2Add( b2Sub( vB, vA ), b2Sub( b2CrossSV( wB, rB ), b2CrossSV( wA, rA ) ) );
			float Cdot = b2Dot( axis, vr );
			float impulse = joint->axialMass * ( joint->motorSpeed - Cdot );
			float oldImpulse = joint->motorImpulse;
			float maxImpulse = h * joint->maxMotorForce;
			joint->motorImpulse = b2ClampFloat( joint->motorImpulse + impulse, -maxImpulse, maxImpulse );
			impulse = joint->motorImpulse - oldImpulse;

			b2Vec2 P = b2MulSV( impulse, axis );
			vA = b2MulSub( vA, mA, P );
			wA -= iA * b2Cross( rA, P );
			vB = b2MulAdd( vB, mB, P );
			wB += iB * b2Cross( rB, P );
		}

		if ( joint->enableLimit )
		{
			// lower
			{
				b2Vec2 vr = b2Add( b2Sub( vB, vA ), b2Sub( b2CrossSV( wB, rB ), b2CrossSV( wA, rA ) ) );
				float cdot = b2Dot( axis, vr );
				float c = length - joint->minLength;
				LimitRow r = MakeLimitRow( c, c * inv_h, useBias, joint->constraintSoftness );

				float impulse = -r.massScale * joint->axialMass * ( cdot + r.bias ) - r.impulseScale * joint->lowerImpulse;
				float newImpulse = b2MaxFloat( 0.0f, joint->lowerImpulse + impulse );
				impulse = newImpulse - joint->lowerImpulse;
				joint->lowerImpulse = newImpulse;

				b2Vec2 P = b2MulSV( impulse, axis );
				vA = b2MulSub( vA, mA, P );
				wA -= iA * b2Cross( rA, P );
				vB = b2MulAdd( vB, mB, P );
				wB += iB * b2Cross( rB, P );
			}

			// upper
			{
				b2Vec2 vr = b2Add( b2Sub( vA, vB ), b2Sub( b2CrossSV( wA, rA ), b2CrossSV( wB, rB ) ) );
				float Cdot = b2Dot( axis, vr );
				float C = joint->maxLength - length;
				LimitRow r = MakeLimitRow( C, C * inv_h, useBias, joint->constraintSoftness );

				float impulse = -r.massScale * joint->axialMass * ( Cdot + r.bias ) - r.impulseScale * joint->upperImpulse;
				float newImpulse = b2MaxFloat( 0.0f, joint->upperImpulse + impulse );
				impulse = newImpulse - joint->upperImpulse;
				joint->upperImpulse = newImpulse;

				b2Vec2 P = b2MulSV( -impulse, axis );
				vA = b2MulSub( vA, mA, P );
				wA -= iA * b2Cross( rA, P );
				vB = b2MulAdd( vB, mB, P );
				wB += iB * b2Cross( rB, P );
			}
		}
	}
	else
	{
		// rigid
		b2Vec2 vr = b2Add( b2Sub( vB, vA ), b2Sub( b2CrossSV( wB, rB ), b2CrossSV( wA, rA ) ) );
		float Cdot = b2Dot( axis, vr );

		float C = length - joint->length;

		float bias = 0.0f;
		float massScale = 1.0f;
		float impulseScale = 0.0f;
		if ( useBias )
		{
			bias = joint->constraintSoftness.biasRate * C;
			massScale = joint->constraintSoftness.massScale;
			impulseScale = joint->constraintSoftness.impulseScale;
		}

		float impulse = -massScale * joint->axialMass * ( Cdot + bias ) - impulseScale * joint->distanceImpulse;
		joint->distanceImpulse += impulse;

		b2Vec2 P = b2MulSV( impulse, axis );
		vA = b2MulSub( vA, mA, P );
		wA -= iA * b2Cross( rA, P );
		vB = b2MulAdd( vB, mB, P );
		wB += iB * b2Cross( rB, P );
	}

	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = vA;
		sA.angularVelocity = wA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = vB;
		sB.angularVelocity = wB;
	}
	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Motor

template <typename Bodies>
B2G_DEVI void SolveMotor( JointSim* joint, const Bodies& B, float h )
{
	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 vA = sA.linearVelocity;
	float wA = sA.angularVelocity;
	b2Vec2 vB = sB.linearVelocity;
	float wB = sB.angularVelocity;

	// angular spring
	if ( joint->maxSpringTorque > 0.0f && joint->angularHertz > 0.0f )
	{
		b2Rot qA = b2MulRot( sA.deltaRotation, joint->frameA.q );
		b2Rot qB = b2MulRot( sB.deltaRotation, joint->frameB.q );
		b2Rot relQ = b2InvMulRot( qA, qB );

		float c = Atan2( relQ.s, relQ.c );
		float bias = joint->angularSpring.biasRate * c;
		float massScale = joint->angularSpring.massScale;
		float impulseScale = joint->angularSpring.impulseScale;

		float cdot = wB - wA;

		float maxImpulse = h * joint->maxSpringTorque;
		float oldImpulse = joint->angularSpringImpulse;
		float impulse = -massScale * joint->angularMass * ( cdot + bias ) - impulseScale * oldImpulse;
		joint->angularSpringImpulse = b2ClampFloat( oldImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint->angularSpringImpulse - oldImpulse;

		wA -= iA * impulse;
		wB += iB * impulse;
	}

	// angular velocity
	if ( joint->maxVelocityTorque > 0.0f )
	{
		float cdot = wB - wA - joint->angularVelocity;
		float impulse = -joint->angularMass * cdot;

		float maxImpulse = h * joint->maxVelocityTorque;
		float oldImpulse = joint->angularVelocityImpulse;
		joint->angularVelocityImpulse = b2ClampFloat( oldImpulse + impulse, -maxImpulse, maxImpulse );
		impulse = joint->angularVelocityImpulse - oldImpulse;

		wA -= iA * impulse;
		wB += iB * impulse;
	}

	b2Vec2 rA = b2RotateVector( sA.deltaRotation, joint->frameA.p );
	b2Vec2 rB = b2RotateVector( sB.deltaRotation, joint->frameB.p );

	// linear spring
	if ( joint->maxSpringForce > 0.0f && joint->linearHertz > 0.0f )
	{
		b2Vec2 dcA = sA.deltaPosition;
		b2Vec2 dcB = sB.deltaPosition;
		b2Vec2 c = b2Add( b2Add( b2Sub( dcB, dcA ), b2Sub( rB, rA ) ), joint->deltaCenter );

		b2Vec2 bias = b2MulSV( joint->linearSpring.biasRate, c );
		float massScale = joint->linearSpring.massScale;
		float impulseScale = joint->linearSpring.impulseScale;

		b2Vec2 cdot = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );
		cdot = b2Add( cdot, bias );

		// the effective mass is refreshed here with the current anchors
		b2Mat22 kl;
		kl.cx.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
		kl.cx.y = -rA.y * rA.x * iA - rB.y * rB.x * iB;
		kl.cy.x = kl.cx.y;
		kl.cy.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
		joint->linearMass = b2GetInverse22( kl );

		b2Vec2 b = b2MulMV( joint->linearMass, cdot );

		b2Vec2 oldImpulse = joint->linearSpringImpulse;
		b2Vec2 impulse = {
			-massScale * b.x - impulseScale * oldImpulse.x,
			-massScale * b.y - impulseScale * oldImpulse.y,
		};

		float maxImpulse = h * joint->maxSpringForce;
		joint->linearSpringImpulse = b2Add( joint->linearSpringImpulse, impulse );

		if ( b2LengthSquared( joint->linearSpringImpulse ) > maxImpulse * maxImpulse )
		{
			joint->linearSpringImpulse = b2Normalize( joint->linearSpringImpulse );
			joint->linearSpringImpulse.x *= maxImpulse;
			joint->linearSpringImpulse.y *= maxImpulse;
		}

		impulse = b2Sub( joint->linearSpringImpulse, oldImpulse );

		vA = b2MulSub( vA, mA, impulse );
		wA -= iA * b2Cross( rA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}

	// linear velocity
	if ( joint->maxVelocityForce > 0.0f )
	{
		b2Vec2 cdot = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );
		cdot = b2Sub( cdot, joint->linearVelocity );
		b2Vec2 b = b2MulMV( joint->linearMass, cdot );
		b2Vec2 impulse = { -b.x, -b.y };

		b2Vec2 oldImpulse = joint->linearVelocityImpulse;
		float maxImpulse = h * joint->maxVelocityForce;
		joint->linearVelocityImpulse = b2Add( joint->linearVelocityImpulse, impulse );

		if ( b2LengthSquared( joint->linearVelocityImpulse ) > maxImpulse * maxImpulse )
		{
			joint->linearVelocityImpulse = b2Normalize( joint->linearVelocityImpulse );
			joint->linearVelocityImpulse.x *= maxImpulse;
			joint->linearVelocityImpulse.y *= maxImpulse;
		}

		impulse = b2Sub( joint->linearVelocityImpulse, oldImpulse );

		vA = b2MulSub( vA, mA, impulse );
		wA -= iA * b2Cross( rA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
		wB += iB * b2Cross( rB, impulse );
	}

	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = vA;
		sA.angularVelocity = wA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = vB;
		sB.angularVelocity = wB;
	}
	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Mover

template <typename Bodies>
B2G_DEVI void SolveMover( JointSim* joint, const Bodies& B, float h )
{
	float mA = joint->invMassA;
	float mB = joint->invMassB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 vA = sA.linearVelocity;
	b2Vec2 vB = sB.linearVelocity;

	if ( joint->maxVelocityForce2.x > 0.0f || joint->maxVelocityForce2.y > 0.0f )
	{
		b2Vec2 cdot = b2Sub( vB, vA );
		cdot = b2Sub( cdot, joint->linearVelocity );
		b2Vec2 b = b2MulSV( joint->pointMass, cdot );
		b2Vec2 impulse = { -b.x, -b.y };

		b2Vec2 oldImpulse = joint->linearVelocityImpulse;
		joint->linearVelocityImpulse = b2Add( joint->linearVelocityImpulse, impulse );
		b2Vec2 maxImpulse = b2MulSV( h, joint->maxVelocityForce2 );

		joint->linearVelocityImpulse = b2Clamp( joint->linearVelocityImpulse, b2Neg( maxImpulse ), maxImpulse );

		impulse = b2Sub( joint->linearVelocityImpulse, oldImpulse );
		vA = b2MulSub( vA, mA, impulse );
		vB = b2MulAdd( vB, mB, impulse );
	}
	else
	{
		joint->linearVelocityImpulse = { 0.0f, 0.0f };
	}

	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = vA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = vB;
	}
	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Pogo

template <typename Bodies>
B2G_DEVI void SolvePogo( JointSim* joint, const Bodies& B, bool useBias, float h )
{
	if ( joint->hertz == 0.0f )
	{
		joint->pogoImpulse = 0.0f;
		return;
	}

	float mA = joint->invMassA;
	float mB = joint->invMassB;
	float iA = joint->invIA;
	float iB = joint->invIB;
	int ia = Bodies::JointA( joint );
	int ib = Bodies::JointB( joint );
	BodyState sA = ToState( B.Load( ia ) );
	BodyState sB = ToState( B.Load( ib ) );

	b2Vec2 vA = sA.linearVelocity;
	float wA = sA.angularVelocity;
	b2Vec2 vB = sB.linearVelocity;
	float wB = sB.angularVelocity;

	b2Vec2 rA = b2RotateVector( sA.deltaRotation, joint->frameA.p );
	b2Vec2 rB = b2RotateVector( sB.deltaRotation, joint->frameB.p );

	float bias = 0.0f;
	if ( useBias )
	{
		// position correction through a spring-damper on the pogo axis (body B's local y axis),
		// so no velocity is fed into the driven body
		b2Vec2 dcA = sA.deltaPosition;
		b2Vec2 dcB = sB.deltaPosition;
		b2Vec2 d = b2Add( b2Add( b2Sub( dcB, dcA ), b2Sub( rB, rA ) ), joint->deltaCenter );

		b2Vec2 pogoAxis = b2RotateVector( joint->frameB.q, b2Vec2{ 0.0f, 1.0f } );
		float c = b2Dot( pogoAxis, d ) - joint->restLength;

		joint->pogoVelocity = b2SpringDamper( joint->hertz, joint->dampingRatio, c, joint->pogoVelocity, h );
		bias = -joint->pogoVelocity;
	}

	b2Vec2 vr = b2Sub( b2Add( vB, b2CrossSV( wB, rB ) ), b2Add( vA, b2CrossSV( wA, rA ) ) );
	float cdot = b2Dot( joint->normal, vr );

	float maxTensionImpulse = h * joint->maxTensionForce;
	float maxCompressionImpulse = h * joint->maxCompressionForce;
	float oldImpulse = joint->pogoImpulse;
	float impulse = -joint->pointMass * ( cdot + bias );
	joint->pogoImpulse = b2ClampFloat( joint->pogoImpulse + impulse, -maxTensionImpulse, maxCompressionImpulse );
	impulse = joint->pogoImpulse - oldImpulse;

	b2Vec2 P = b2MulSV( impulse, joint->normal );
	vA = b2MulSub( vA, mA, P );
	wA -= iA * b2Cross( rA, P );
	vB = b2MulAdd( vB, mB, P );
	wB += iB * b2Cross( rB, P );

	if ( sA.flags & kDynamicFlag )
	{
		sA.linearVelocity = vA;
		sA.angularVelocity = wA;
	}
	if ( sB.flags & kDynamicFlag )
	{
		sB.linearVelocity = vB;
		sB.angularVelocity = wB;
	}
	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

// ---------------------------------------------------------------------------------------------
// Joint events (reference src/joint.c:993-1072 and the threshold test at src/joint.c:1663-1675):
// the constraint force and torque after a biased solve, compared with the joint's thresholds.

B2G_DEVI bool JointReactionReached( const JointSim* j, float inv_h )
{
	float linearImpulse = 0.0f;
	float angularImpulse = 0.0f;
	switch ( j->type )
	{
		case b2_distanceJoint:
			linearImpulse = b2AbsFloat( j->distanceImpulse + j->lowerImpulse - j->upperImpulse + j->motorImpulse );
			break;
		case b2_motorJoint:
			linearImpulse = b2Length( b2Add( j->linearVelocityImpulse, j->linearSpringImpulse ) );
			angularImpulse = b2AbsFloat( j->angularVelocityImpulse + j->angularSpringImpulse );
			break;
		case b2_moverJoint:
			linearImpulse = b2Length( j->linearVelocityImpulse );
			break;
		case b2_pogoJoint:
			linearImpulse = b2AbsFloat( j->pogoImpulse );
			break;
		case b2_prismaticJoint:
		{
			float perp = j->pointImpulse.x;
			float axial = j->motorImpulse + j->lowerImpulse - j->upperImpulse;
			linearImpulse = sqrtf( perp * perp + axial * axial );
			angularImpulse = b2AbsFloat( j->pointImpulse.y );
			break;
		}
		case b2_revoluteJoint:
			linearImpulse = b2Length( j->linearImpulse );
			angularImpulse = b2AbsFloat( j->motorImpulse + j->lowerImpulse - j->upperImpulse );
			break;
		case b2_weldJoint:
			linearImpulse = b2Length( j->linearImpulse );
			angularImpulse = b2AbsFloat( j->angularImpulse );
			break;
		case b2_wheelJoint:
		{
			float perp = j->perpImpulse;
			float axial = j->springImpulse + j->lowerImpulse - j->upperImpulse;
			linearImpulse = sqrtf( perp * perp + axial * axial );
			angularImpulse = b2AbsFloat( j->motorImpulse );
			break;
		}
		default:
			break;
	}
	float force = linearImpulse * inv_h;
	float torque = angularImpulse * inv_h;
	return force >= j->forceThreshold || torque >= j->torqueThreshold;
}

// ---------------------------------------------------------------------------------------------
// Dispatch (src/joint.c:1406-1532)

template <typename Bodies>
B2G_DEVI void SolveJoint( JointSim* joint, const Bodies& B, bool useBias, float h, float inv_h )
{
	switch ( joint->type )
	{
		case b2_revoluteJoint:
			SolveRevolute( joint, B, useBias, h, inv_h );
			break;
		case b2_weldJoint:
			SolveWeld( joint, B, useBias );
			break;
		case b2_prismaticJoint:
			SolvePrismatic( joint, B, useBias, h, inv_h );
			break;
		case b2_wheelJoint:
			SolveWheel( joint, B, useBias, h, inv_h );
			break;
		case b2_distanceJoint:
			SolveDistance( joint, B, useBias, h, inv_h );
			break;
		case b2_motorJoint:
			SolveMotor( joint, B, h );
			break;
		case b2_moverJoint:
			SolveMover( joint, B, h );
			break;
		case b2_pogoJoint:
			SolvePogo( joint, B, useBias, h );
			break;
		default:
			break;
	}
}

} // namespace b2g
