#pragma once
#include "../../include/dwba_hip.h"
