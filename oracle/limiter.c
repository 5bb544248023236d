#include "oracle_internal.h"
