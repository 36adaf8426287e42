/* pq_host.h -- internal declarations shared by the host-side C files. */
#ifndef PQ_HOST_H
#define PQ_HOST_H

#include "../../../include/pasqual_amd.h"

#endif
