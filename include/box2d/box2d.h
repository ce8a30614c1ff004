// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the declarations below follow its public headers)
// SPDX-License-Identifier: MIT
// box2d_amd public API: the Box2D world, body, shape and joint calls.
//
// Drop-in counterpart of the reference's include/box2d/box2d.h for the world step path. Each
// declaration keeps the reference's name, signature and meaning; the file:line after each group
// names the reference declaration it replaces. b2World_Step runs the step on the GPU.
#pragma once

#include "base.h"
#include "collision.h"
#include "id.h"
#include "types.h"

/// World (reference: include/box2d/box2d.h:34-228)
B2_API b2WorldId b2CreateWorld( const b2WorldDef* def );
B2_API void b2DestroyWorld( b2WorldId worldId );
B2_API bool b2World_IsValid( b2WorldId id );
B2_API void b2World_Step( b2WorldId worldId, float timeStep, int subStepCount );
B2_API b2BodyEvents b2World_GetBodyEvents( b2WorldId worldId );
B2_API b2ContactEvents b2World_GetContactEvents( b2WorldId worldId );
/// Joints whose force or torque reached the joint's threshold during the last step (reference:
/// include/box2d/box2d.h:64).
B2_API b2JointEvents b2World_GetJointEvents( b2WorldId worldId );
/// Sensor begin and end overlap events of the last step (reference: include/box2d/box2d.h:58).
B2_API b2SensorEvents b2World_GetSensorEvents( b2WorldId worldId );
B2_API void b2World_EnableSleeping( b2WorldId worldId, bool flag );
B2_API bool b2World_IsSleepingEnabled( b2WorldId worldId );
B2_API void b2World_EnableContinuous( b2WorldId worldId, bool flag );
B2_API bool b2World_IsContinuousEnabled( b2WorldId worldId );
B2_API void b2World_EnableWarmStarting( b2WorldId worldId, bool flag );
B2_API bool b2World_IsWarmStartingEnabled( b2WorldId worldId );
B2_API void b2World_SetGravity( b2WorldId worldId, b2Vec2 gravity );
B2_API b2Vec2 b2World_GetGravity( b2WorldId worldId );
B2_API void b2World_SetContactTuning( b2WorldId worldId, float hertz, float dampingRatio, float pushSpeed );
B2_API void b2World_SetMaximumLinearSpeed( b2WorldId worldId, float maximumLinearSpeed );
B2_API float b2World_GetMaximumLinearSpeed( b2WorldId worldId );
B2_API int b2World_GetAwakeBodyCount( b2WorldId worldId );
B2_API b2Profile b2World_GetProfile( b2WorldId worldId );
B2_API b2Counters b2World_GetCounters( b2WorldId worldId );
B2_API void b2World_SetUserData( b2WorldId worldId, void* userData );
B2_API void* b2World_GetUserData( b2WorldId worldId );
/// World settings and callbacks (reference: include/box2d/box2d.h:127-144, 209-213). The callbacks
/// run on the host: the custom filter where the pair update filters candidates, the mixing
/// callbacks when a contact is created and whenever a full narrowphase update follows a change of
/// callback or shape material.
B2_API void b2World_SetRestitutionThreshold( b2WorldId worldId, float value );
B2_API float b2World_GetRestitutionThreshold( b2WorldId worldId );
B2_API void b2World_SetHitEventThreshold( b2WorldId worldId, float value );
B2_API float b2World_GetHitEventThreshold( b2WorldId worldId );
B2_API void b2World_SetCustomFilterCallback( b2WorldId worldId, b2CustomFilterFcn* fcn, void* context );
/// Pre-solve callback, run on the host for the contacts of shapes with pre-solve events enabled
/// (their narrowphase runs on the host around it). A pre-continuous callback is accepted, but a
/// step with it set and a shape with pre-solve events enabled stops with an error, because
/// continuous collision runs on the GPU (reference: include/box2d/box2d.h:146-148).
B2_API void b2World_SetPreSolveCallback( b2WorldId worldId, b2PreSolveFcn* preSolveFcn, b2PreContinuousFcn* preContinuousFcn,
										 void* context );
B2_API void b2World_SetFrictionCallback( b2WorldId worldId, b2FrictionCallback* callback );
B2_API void b2World_SetRestitutionCallback( b2WorldId worldId, b2RestitutionCallback* callback );
/// Deterministic hash of the simulation state: body transforms and awake velocities, contact and
/// joint impulses, and the set/color/local indices and id pools that drive the solve (reference:
/// include/box2d/box2d.h:230-235, computed as src/world_snapshot.c:1025-1187 does).
B2_API uint64_t b2World_GetStateHash( b2WorldId worldId );
/// Snapshots of a world's simulation state between steps (reference: include/box2d/box2d.h:287-337).
/// The image layout is this library's own; it round-trips through b2World_Restore and
/// b2CreateWorldFromSnapshot, and the restored world steps exactly as the original would. User data
/// and callbacks are not stored. Restore returns false and leaves the world unchanged for an image
/// it rejects (magic, version, layout or checksum).
B2_API int b2World_Snapshot( b2WorldId worldId, uint8_t* image, int capacity );
B2_API bool b2World_Restore( b2WorldId worldId, const uint8_t* image, int size );
B2_API b2WorldId b2CreateWorldFromSnapshot( const uint8_t* image, int size, int workerCount );

/// Bodies (reference: include/box2d/box2d.h:355-645)
B2_API b2BodyId b2CreateBody( b2WorldId worldId, const b2BodyDef* def );
B2_API bool b2Body_IsValid( b2BodyId id );
B2_API b2BodyType b2Body_GetType( b2BodyId bodyId );
B2_API b2Pos b2Body_GetPosition( b2BodyId bodyId );
B2_API b2Rot b2Body_GetRotation( b2BodyId bodyId );
B2_API b2WorldTransform b2Body_GetTransform( b2BodyId bodyId );
B2_API b2Pos b2Body_GetWorldCenterOfMass( b2BodyId bodyId );
B2_API b2Vec2 b2Body_GetLocalCenterOfMass( b2BodyId bodyId );
B2_API b2Vec2 b2Body_GetLinearVelocity( b2BodyId bodyId );
B2_API float b2Body_GetAngularVelocity( b2BodyId bodyId );
B2_API void b2Body_SetLinearVelocity( b2BodyId bodyId, b2Vec2 linearVelocity );
B2_API void b2Body_SetAngularVelocity( b2BodyId bodyId, float angularVelocity );
B2_API float b2Body_GetMass( b2BodyId bodyId );
B2_API float b2Body_GetRotationalInertia( b2BodyId bodyId );
B2_API bool b2Body_IsAwake( b2BodyId bodyId );
B2_API int b2Body_GetShapeCount( b2BodyId bodyId );
B2_API void b2Body_SetUserData( b2BodyId bodyId, void* userData );
B2_API void* b2Body_GetUserData( b2BodyId bodyId );
B2_API b2WorldId b2Body_GetWorld( b2BodyId bodyId );
B2_API void b2DestroyBody( b2BodyId bodyId );
B2_API void b2Body_SetTransform( b2BodyId bodyId, b2Pos position, b2Rot rotation );
B2_API b2Vec2 b2Body_GetLocalPoint( b2BodyId bodyId, b2Pos worldPoint );
B2_API b2Pos b2Body_GetWorldPoint( b2BodyId bodyId, b2Vec2 localPoint );
B2_API b2Vec2 b2Body_GetLocalVector( b2BodyId bodyId, b2Vec2 worldVector );
B2_API b2Vec2 b2Body_GetWorldVector( b2BodyId bodyId, b2Vec2 localVector );
B2_API void b2Body_SetTargetTransform( b2BodyId bodyId, b2WorldTransform target, float timeStep, bool wake );
B2_API void b2Body_ApplyForce( b2BodyId bodyId, b2Vec2 force, b2Pos point, bool wake );
B2_API void b2Body_ApplyForceToCenter( b2BodyId bodyId, b2Vec2 force, bool wake );
B2_API void b2Body_ApplyTorque( b2BodyId bodyId, float torque, bool wake );
B2_API void b2Body_ApplyLinearImpulse( b2BodyId bodyId, b2Vec2 impulse, b2Pos point, bool wake );
B2_API void b2Body_ApplyLinearImpulseToCenter( b2BodyId bodyId, b2Vec2 impulse, bool wake );
B2_API void b2Body_ApplyAngularImpulse( b2BodyId bodyId, float impulse, bool wake );
B2_API void b2Body_ApplyMassFromShapes( b2BodyId bodyId );
B2_API int b2Body_GetShapes( b2BodyId bodyId, b2ShapeId* shapeArray, int capacity );
B2_API int b2Body_GetJointCount( b2BodyId bodyId );
B2_API int b2Body_GetJoints( b2BodyId bodyId, b2JointId* jointArray, int capacity );
/// Damping, gravity scale and sleep (reference: include/box2d/box2d.h:526-582)
B2_API void b2Body_SetLinearDamping( b2BodyId bodyId, float linearDamping );
B2_API float b2Body_GetLinearDamping( b2BodyId bodyId );
B2_API void b2Body_SetAngularDamping( b2BodyId bodyId, float angularDamping );
B2_API float b2Body_GetAngularDamping( b2BodyId bodyId );
B2_API void b2Body_SetGravityScale( b2BodyId bodyId, float gravityScale );
B2_API float b2Body_GetGravityScale( b2BodyId bodyId );
B2_API void b2Body_SetAwake( b2BodyId bodyId, bool awake );
B2_API void b2Body_WakeTouching( b2BodyId bodyId );
B2_API void b2Body_EnableSleep( b2BodyId bodyId, bool enableSleep );
B2_API bool b2Body_IsSleepEnabled( b2BodyId bodyId );
B2_API void b2Body_SetSleepThreshold( b2BodyId bodyId, float sleepThreshold );
B2_API float b2Body_GetSleepThreshold( b2BodyId bodyId );
B2_API bool b2Body_IsEnabled( b2BodyId bodyId );
/// Bullets: continuous collision against moving bodies too (reference: include/box2d/box2d.h:592-596)
B2_API void b2Body_SetBullet( b2BodyId bodyId, bool flag );
B2_API bool b2Body_IsBullet( b2BodyId bodyId );
/// Contact data (reference: include/box2d/box2d.h:633-641, 833-841, 1519-1525)
B2_API int b2Body_GetContactCapacity( b2BodyId bodyId );
B2_API int b2Body_GetContactData( b2BodyId bodyId, b2ContactData* contactData, int capacity );
B2_API int b2Shape_GetContactCapacity( b2ShapeId shapeId );
B2_API int b2Shape_GetContactData( b2ShapeId shapeId, b2ContactData* contactData, int capacity );
B2_API bool b2Contact_IsValid( b2ContactId id );
B2_API b2ContactData b2Contact_GetData( b2ContactId contactId );
/// Type changes and enabling (reference: include/box2d/box2d.h:366-372, 576-582)
B2_API void b2Body_SetType( b2BodyId bodyId, b2BodyType type );
B2_API void b2Body_Disable( b2BodyId bodyId );
B2_API void b2Body_Enable( b2BodyId bodyId );

/// Shapes (reference: include/box2d/box2d.h:659-908)
B2_API b2ShapeId b2CreateCircleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Circle* circle );
B2_API b2ShapeId b2CreateCapsuleShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Capsule* capsule );
B2_API b2ShapeId b2CreatePolygonShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Polygon* polygon );
B2_API b2ShapeId b2CreateSegmentShape( b2BodyId bodyId, const b2ShapeDef* def, const b2Segment* segment );
B2_API bool b2Shape_IsValid( b2ShapeId id );
B2_API b2ShapeType b2Shape_GetType( b2ShapeId shapeId );
B2_API b2BodyId b2Shape_GetBody( b2ShapeId shapeId );
B2_API b2AABB b2Shape_GetAABB( b2ShapeId shapeId );
B2_API void b2Shape_EnableSensorEvents( b2ShapeId shapeId, bool flag );
B2_API bool b2Shape_AreSensorEventsEnabled( b2ShapeId shapeId );
/// The shapes a sensor overlapped at the end of the last step, by shape id (reference:
/// include/box2d/box2d.h:846-856).
B2_API int b2Shape_GetSensorCapacity( b2ShapeId shapeId );
B2_API int b2Shape_GetSensorData( b2ShapeId shapeId, b2ShapeId* visitorIds, int capacity );
/// Pre-solve events of a shape; contacts created afterwards follow the flag (reference:
/// include/box2d/box2d.h:770-774).
B2_API void b2Shape_EnablePreSolveEvents( b2ShapeId shapeId, bool flag );
B2_API bool b2Shape_ArePreSolveEventsEnabled( b2ShapeId shapeId );
B2_API b2Circle b2Shape_GetCircle( b2ShapeId shapeId );
B2_API b2Segment b2Shape_GetSegment( b2ShapeId shapeId );
B2_API b2Capsule b2Shape_GetCapsule( b2ShapeId shapeId );
B2_API b2Polygon b2Shape_GetPolygon( b2ShapeId shapeId );
B2_API void b2Shape_SetCircle( b2ShapeId shapeId, const b2Circle* circle );
B2_API void b2Shape_SetCapsule( b2ShapeId shapeId, const b2Capsule* capsule );
B2_API void b2Shape_SetSegment( b2ShapeId shapeId, const b2Segment* segment );
B2_API void b2Shape_SetPolygon( b2ShapeId shapeId, const b2Polygon* polygon );
/// Shape destruction, materials, density, filter and user data (reference:
/// include/box2d/box2d.h:683-760)
B2_API void b2DestroyShape( b2ShapeId shapeId, bool updateBodyMass );
B2_API b2WorldId b2Shape_GetWorld( b2ShapeId shapeId );
B2_API bool b2Shape_IsSensor( b2ShapeId shapeId );
B2_API void b2Shape_SetUserData( b2ShapeId shapeId, void* userData );
B2_API void* b2Shape_GetUserData( b2ShapeId shapeId );
B2_API void b2Shape_SetDensity( b2ShapeId shapeId, float density, bool updateBodyMass );
B2_API float b2Shape_GetDensity( b2ShapeId shapeId );
B2_API void b2Shape_SetFriction( b2ShapeId shapeId, float friction );
B2_API float b2Shape_GetFriction( b2ShapeId shapeId );
B2_API void b2Shape_SetRestitution( b2ShapeId shapeId, float restitution );
B2_API float b2Shape_GetRestitution( b2ShapeId shapeId );
B2_API void b2Shape_SetUserMaterial( b2ShapeId shapeId, uint64_t material );
B2_API uint64_t b2Shape_GetUserMaterial( b2ShapeId shapeId );
B2_API void b2Shape_SetSurfaceMaterial( b2ShapeId shapeId, const b2SurfaceMaterial* surfaceMaterial );
B2_API b2SurfaceMaterial b2Shape_GetSurfaceMaterial( b2ShapeId shapeId );
B2_API b2Filter b2Shape_GetFilter( b2ShapeId shapeId );
B2_API void b2Shape_SetFilter( b2ShapeId shapeId, b2Filter filter );
B2_API b2ChainId b2CreateChain( b2BodyId bodyId, const b2ChainDef* def );

/// Joints (reference: include/box2d/box2d.h:919-1120)
B2_API b2JointId b2CreateRevoluteJoint( b2WorldId worldId, const b2RevoluteJointDef* def );
B2_API b2JointId b2CreateDistanceJoint( b2WorldId worldId, const b2DistanceJointDef* def );
B2_API b2JointId b2CreateFilterJoint( b2WorldId worldId, const b2FilterJointDef* def );
B2_API b2JointId b2CreateMotorJoint( b2WorldId worldId, const b2MotorJointDef* def );
B2_API b2JointId b2CreateMoverJoint( b2WorldId worldId, const b2MoverJointDef* def );
B2_API b2JointId b2CreatePogoJoint( b2WorldId worldId, const b2PogoJointDef* def );
B2_API b2JointId b2CreatePrismaticJoint( b2WorldId worldId, const b2PrismaticJointDef* def );
B2_API b2JointId b2CreateWeldJoint( b2WorldId worldId, const b2WeldJointDef* def );
B2_API b2JointId b2CreateWheelJoint( b2WorldId worldId, const b2WheelJointDef* def );
B2_API bool b2Joint_IsValid( b2JointId id );
B2_API b2JointType b2Joint_GetType( b2JointId jointId );
B2_API b2BodyId b2Joint_GetBodyA( b2JointId jointId );
B2_API b2BodyId b2Joint_GetBodyB( b2JointId jointId );
B2_API void b2DestroyJoint( b2JointId jointId );
B2_API void b2Joint_SetLocalFrameA( b2JointId jointId, b2Transform localFrame );
B2_API b2Transform b2Joint_GetLocalFrameA( b2JointId jointId );
B2_API void b2Joint_SetLocalFrameB( b2JointId jointId, b2Transform localFrame );
B2_API b2Transform b2Joint_GetLocalFrameB( b2JointId jointId );
/// Joint user data and event thresholds (reference: include/box2d/box2d.h:955-994).
B2_API void b2Joint_SetUserData( b2JointId jointId, void* userData );
B2_API void* b2Joint_GetUserData( b2JointId jointId );
B2_API void b2Joint_SetForceThreshold( b2JointId jointId, float threshold );
B2_API float b2Joint_GetForceThreshold( b2JointId jointId );
B2_API void b2Joint_SetTorqueThreshold( b2JointId jointId, float threshold );
B2_API float b2Joint_GetTorqueThreshold( b2JointId jointId );

/// Revolute joint (reference: include/box2d/box2d.h:1336-1398)
B2_API void b2RevoluteJoint_EnableSpring( b2JointId jointId, bool enableSpring );
B2_API bool b2RevoluteJoint_IsSpringEnabled( b2JointId jointId );
B2_API void b2RevoluteJoint_SetSpringHertz( b2JointId jointId, float hertz );
B2_API float b2RevoluteJoint_GetSpringHertz( b2JointId jointId );
B2_API void b2RevoluteJoint_SetSpringDampingRatio( b2JointId jointId, float dampingRatio );
B2_API float b2RevoluteJoint_GetSpringDampingRatio( b2JointId jointId );
B2_API void b2RevoluteJoint_SetTargetAngle( b2JointId jointId, float angle );
B2_API float b2RevoluteJoint_GetTargetAngle( b2JointId jointId );
B2_API float b2RevoluteJoint_GetAngle( b2JointId jointId );
B2_API void b2RevoluteJoint_EnableLimit( b2JointId jointId, bool enableLimit );
B2_API bool b2RevoluteJoint_IsLimitEnabled( b2JointId jointId );
B2_API float b2RevoluteJoint_GetLowerLimit( b2JointId jointId );
B2_API float b2RevoluteJoint_GetUpperLimit( b2JointId jointId );
B2_API void b2RevoluteJoint_SetLimits( b2JointId jointId, float lower, float upper );
B2_API void b2RevoluteJoint_EnableMotor( b2JointId jointId, bool enableMotor );
B2_API bool b2RevoluteJoint_IsMotorEnabled( b2JointId jointId );
B2_API void b2RevoluteJoint_SetMotorSpeed( b2JointId jointId, float motorSpeed );
B2_API float b2RevoluteJoint_GetMotorSpeed( b2JointId jointId );
B2_API float b2RevoluteJoint_GetMotorTorque( b2JointId jointId );
B2_API void b2RevoluteJoint_SetMaxMotorTorque( b2JointId jointId, float torque );
B2_API float b2RevoluteJoint_GetMaxMotorTorque( b2JointId jointId );
