// Internal include: the public C ABI plus the synthetic-scenario generator.
#pragma once

#include "../../include/gcomcot_hip.h"
#include "gcomcot_synth.h"
