/* placeholder, replaced by the adjoint restatement */
