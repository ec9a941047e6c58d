"""Host helpers mirroring the reference's helpers package (src/helpers/)."""
