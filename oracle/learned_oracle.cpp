// Learned-weight BP oracle (filled in with the learned decoder).
