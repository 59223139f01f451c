#pragma once
#include "../../include/uhs_dsl.h"
