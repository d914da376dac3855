// Boundary.hh -- block edges, ghost-layer states and boundary types, with the
// reference's enumerator names and order (src/types/Boundary.hh:5-28), so
// that values cross the C ABI unchanged.
#ifndef SWE_HIP_BOUNDARY_HH
#define SWE_HIP_BOUNDARY_HH

typedef enum Boundary { BND_LEFT, BND_RIGHT, BND_BOTTOM, BND_TOP } Boundary;

typedef enum GhostlayerState {
    GL_DONE,     // the neighbour has reached the end of the simulation
    GL_UNVALID,  // the neighbour is behind: wait
    GL_INTER,    // the neighbour is ahead: ghost values interpolated
    GL_SYNC,     // the neighbour is at the same time: ghost values copied
    GL_NEXT      // expecting the neighbour's next time level
} GhostlayerState;

typedef enum BoundaryType { OUTFLOW, WALL, INFLOW, CONNECT, CONNECT_WITHIN_RANK, PASSIVE } BoundaryType;

#endif
