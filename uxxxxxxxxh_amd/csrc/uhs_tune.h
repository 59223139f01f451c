#pragma once
#include "../../include/uhs_tune.h"
