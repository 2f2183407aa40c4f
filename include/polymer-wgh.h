/*
 * polymer-wgh.h -- the weighted edgeMap interface (polymer-wgh.h:888-1511 of
 * the reference: wghGraph, functors update(s, d, w) / updateAtomic(s, d, w)).
 * On the GPU one header carries both forms: edgeMap and its variants take a
 * wghGraph<vertex> and pass each edge's int32 weight to the functor.
 */
#ifndef POLYMER_WGH_HIP_H
#define POLYMER_WGH_HIP_H
#include "polymer.h"
#endif
