"""ORACLE — test infrastructure (CPU restatement of the reference path); never imported by sv_raster_amd."""
