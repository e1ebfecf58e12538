"""CPU oracle of the SPH step -- test infrastructure (the checker), never
imported by the product package fluidgrid_amd."""
