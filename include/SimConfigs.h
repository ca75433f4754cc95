// Compile-time configuration (drop-in for the reference's include/SimConfigs.h:1-7).
#pragma once

// When defined, ToucanSimulator::printProfile() is available. The MI355X
// runtime always collects its launch profile, so the flag only exposes the
// reference's method name.
// #define ENABLE_SIM_PROFILE

#define PROFILE_START_CYCLE 100
#define PROFILE_COLLECT_CYCLE 100
