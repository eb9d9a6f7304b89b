/* freeze-out hypersurface restatement: filled in below */
