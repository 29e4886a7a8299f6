// Stands in for the pmbb-ivqm header of the same name: the MI355X drop-in processors.
#pragma once
#include "ivqm_pmbb_shim.h"
